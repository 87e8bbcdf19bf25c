"""CPU restatement of the device's coarsest-level inverse (gj_init_kernel,
gj_step_kernel, gj_mirror_kernel in topoptlab_amd/csrc/tl_mg.hip): blocked
Gauss-Jordan on the lower 64x64 tiles of a symmetric positive definite
matrix, with the row panel of the next pivot block handed from step to step.

For an SPD matrix every Gauss-Jordan intermediate satisfies
A[i][j] = s(i) s(j) A[j][i] with s = -1 on the indices already swept, so the
kernels keep the lower tiles only: the column panel C[i] = A[i][K] is
s(i) R[K][i] of the saved row panel, and the next row panel is gathered from
the pivot tile row (direct) and from the tiles below the pivot tile
(transposed).  The restatement follows the kernels' tile order, buffers and
index ranges; it checks that the result is the inverse, that every buffer
access is in range, and that every entry of the next row panel is written
exactly once.
"""
import numpy as np
import pytest

GJB, T = 32, 64


def tri_unrank(b):
    i = int((np.sqrt(8.0 * b + 1.0) - 1.0) * 0.5)
    while (i + 1) * (i + 2) // 2 <= b:
        i += 1
    while i * (i + 1) // 2 > b:
        i -= 1
    return i, b - i * (i + 1) // 2


def sweep_inverse(A0):
    n = A0.shape[0]
    assert n % GJB == 0
    nt = (n + T - 1) // T
    nbk = n // GJB
    A = A0.copy()
    # init: row panel of block 0 and the inverse of its pivot block
    rowp = A[:GJB, :].copy()
    pinv = np.linalg.inv(A[:GJB, :GJB])
    ntiles = nt * (nt + 1) // 2
    for k in range(nbk):
        kb, kn = k * GJB, (k + 1) * GJB
        has_next = k + 1 < nbk
        tn = kn // T if has_next else None
        owner = tn * (tn + 1) // 2 + tn if has_next else None
        rowp_next = np.full((GJB, n), np.nan)
        written = np.zeros((GJB, n), dtype=int)
        piv = None
        seen = set()
        for hw in range(ntiles):
            lg = hw
            if owner is not None:
                if hw == 0:
                    lg = owner
                elif hw == owner:
                    lg = 0
            I, J = tri_unrank(lg)
            assert 0 <= J <= I < nt
            seen.add((I, J))
            i0, j0 = I * T, J * T
            rows = np.arange(i0, min(i0 + T, n))
            cols = np.arange(j0, min(j0 + T, n))
            aold = A[np.ix_(rows, cols)].copy()
            sig = np.where(rows < kb, -1.0, 1.0)
            C = sig[:, None] * rowp[:, rows].T          # C[i][t] = s(i) R[t][i]
            jin = (cols >= kb) & (cols < kb + GJB)
            X = pinv @ rowp[:, cols]
            X[:, jin] = pinv[:, cols[jin] - kb]
            S = C @ X
            irow = (rows >= kb) & (rows < kb + GJB)
            v = aold - S
            v[:, jin] = -S[:, jin]
            v[irow, :] = X[rows[irow] - kb, :]
            A[np.ix_(rows, cols)] = v
            if has_next:
                inext = (rows >= kn) & (rows < kn + GJB)
                jnext = (cols >= kn) & (cols < kn + GJB)
                # (a) rows of the next pivot block, direct
                for a, i in enumerate(rows):
                    if inext[a]:
                        rowp_next[i - kn, cols] = v[a, :]
                        written[i - kn, cols] += 1
                # (b) tiles below the pivot tile: columns of the block, transposed
                if I > tn:
                    for c, j in enumerate(cols):
                        if jnext[c]:
                            assert np.all(rows >= kn + GJB)  # unswept rows: sign +1
                            rowp_next[j - kn, rows] = v[:, c]
                            written[j - kn, rows] += 1
                if lg == owner:
                    piv = v[np.ix_(inext, jnext)].copy()
        assert len(seen) == ntiles
        if has_next:
            assert np.all(written == 1), "each next row-panel entry written once"
            rowp = rowp_next
            pinv = np.linalg.inv(piv)
    # mirror the lower tiles into the upper ones
    for I in range(nt):
        for J in range(I + 1, nt):
            r = slice(I * T, min(I * T + T, n))
            c = slice(J * T, min(J * T + T, n))
            A[r, c] = A[c, r].T
    return A


@pytest.mark.parametrize("n", [32, 64, 96, 160, 256, 288])
def test_sweep_restatement_is_the_inverse(n):
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n)
    Ainv = sweep_inverse(A)
    assert np.abs(Ainv @ A - np.eye(n)).max() < 1e-10


def test_intermediates_keep_the_signed_symmetry():
    n = 128
    rng = np.random.default_rng(1)
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n)
    # full Gauss-Jordan on blocks of 32, checking A[i][j] = s(i)s(j)A[j][i]
    M = A.copy()
    for k in range(n // GJB):
        K = slice(k * GJB, (k + 1) * GJB)
        P = np.linalg.inv(M[K, K])
        R = P @ M[K, :]
        C = M[:, K].copy()
        M = M - C @ R
        M[K, :] = R
        M[:, K] = -C @ P
        M[K, K] = P
        s = np.where(np.arange(n) < (k + 1) * GJB, -1.0, 1.0)
        assert np.allclose(M, (s[:, None] * s[None, :]) * M.T, rtol=1e-12, atol=1e-12)
    assert np.allclose(M, np.linalg.inv(A))


def test_tile_unrank_covers_the_lower_triangle():
    for nt in (1, 2, 3, 30, 47):
        got = {tri_unrank(b) for b in range(nt * (nt + 1) // 2)}
        assert got == {(i, j) for i in range(nt) for j in range(i + 1)}
