"""numpy restatement of the library's Jacobian-block preconditioners (test
infrastructure): block Jacobi, j-line block-tridiagonal, symmetric
multicolour j-line Gauss-Seidel, five-colour, natural-order and chunked line
block ILU(0), built from a dense Jacobian."""
import numpy as np

SLOTS = [(0, 0), (0, -1), (0, 1), (-1, 0), (1, 0), (-1, -1), (1, -1), (-1, 1), (1, 1), (0, -2), (0, 2), (-2, 0), (2, 0)]


def blocks_from_dense(J, nic, njc, nv):
    """B[i][j][s] = J[(i,j), (i,j)+SLOTS[s]] (nv x nv), zero outside the grid."""
    Jr = J.reshape(nic, njc, nv, nic, njc, nv)
    B = np.zeros((nic, njc, len(SLOTS), nv, nv))
    for i in range(nic):
        for j in range(njc):
            for s, (di, dj) in enumerate(SLOTS):
                ii, jj = i + di, j + dj
                if 0 <= ii < nic and 0 <= jj < njc:
                    B[i, j, s] = Jr[i, j, :, ii, jj, :]
    return B


def line_solve(B, idt, i, t, jacobi=False):
    """Solve the i-th j-line (I/dt - J) restricted to slots 0,1,2 (or 0)."""
    njc, nv = t.shape
    A = np.zeros((njc * nv, njc * nv))
    for j in range(njc):
        A[j * nv:(j + 1) * nv, j * nv:(j + 1) * nv] = np.eye(nv) * idt[i, j] - B[i, j, 0]
        if not jacobi:
            if j > 0:
                A[j * nv:(j + 1) * nv, (j - 1) * nv:j * nv] = -B[i, j, 1]
            if j < njc - 1:
                A[j * nv:(j + 1) * nv, (j + 1) * nv:(j + 2) * nv] = -B[i, j, 2]
    return np.linalg.solve(A, t.reshape(-1)).reshape(njc, nv)


def apply(B, idt, v, precond, sweeps=1, width=16):
    v = np.asarray(v, dtype=np.float64)
    nic, njc, nv = v.shape
    if precond == 0:
        return v.copy()
    if precond == 6:
        return ilu_chunked(B, idt, v, width)
    if precond == 4:
        return ilu_mc5(B, idt, v)
    if precond == 5:
        return ilu_natural(B, idt, v)
    if precond in (1, 2):
        return np.stack([line_solve(B, idt, i, v[i], jacobi=(precond == 1)) for i in range(nic)])
    z = np.zeros_like(v)
    for _ in range(sweeps):
        for colour in (0, 1, 2, 2, 1, 0):
            for i in range(colour, nic, 3):
                t = v[i].copy()
                for j in range(njc):
                    for s in range(3, len(SLOTS)):
                        di, dj = SLOTS[s]
                        ii, jj = i + di, j + dj
                        if 0 <= ii < nic and 0 <= jj < njc:
                            t[j] += B[i, j, s] @ z[ii, jj]
                z[i] = line_solve(B, idt, i, t)
    return z


def colour5(i, j):
    return (i + 3 * j) % 5


def ilu_mc5(B, idt, v):
    """Five-colour block ILU(0) of I/dt - J on the 13-point stencil applied to
    v: factor rows colour by colour (lower neighbours eliminated in colour
    order), then forward/backward solves by colour."""
    v = np.asarray(v, dtype=np.float64)
    nic, njc, nv = v.shape
    A = {}
    for i in range(nic):
        for j in range(njc):
            for s, (di, dj) in enumerate(SLOTS):
                ii, jj = i + di, j + dj
                if 0 <= ii < nic and 0 <= jj < njc:
                    blk = -B[i, j, s].copy()
                    if s == 0:
                        blk += np.eye(nv) * idt[i, j]
                    A[(i, j, ii, jj)] = blk
    Dinv = {}
    for c in range(5):
        for i in range(nic):
            for j in range(njc):
                if colour5(i, j) != c:
                    continue
                for cn in range(c):
                    for (di, dj) in SLOTS[1:]:
                        ni, nj = i + di, j + dj
                        if not (0 <= ni < nic and 0 <= nj < njc) or colour5(ni, nj) != cn:
                            continue
                        L = A[(i, j, ni, nj)] @ Dinv[(ni, nj)]
                        A[(i, j, ni, nj)] = L
                        for (ti, tj) in SLOTS[1:]:
                            ji, jj = ni + ti, nj + tj
                            if not (0 <= ji < nic and 0 <= jj < njc) or colour5(ji, jj) <= cn:
                                continue
                            if (ji - i, jj - j) not in SLOTS:
                                continue
                            A[(i, j, ji, jj)] = A[(i, j, ji, jj)] - L @ A[(ni, nj, ji, jj)]
                Dinv[(i, j)] = np.linalg.inv(A[(i, j, i, j)])
    z = np.zeros_like(v)
    for c in range(5):
        for i in range(nic):
            for j in range(njc):
                if colour5(i, j) != c:
                    continue
                acc = v[i, j].copy()
                for (di, dj) in SLOTS[1:]:
                    ni, nj = i + di, j + dj
                    if 0 <= ni < nic and 0 <= nj < njc and colour5(ni, nj) < c:
                        acc -= A[(i, j, ni, nj)] @ z[ni, nj]
                z[i, j] = acc
    for c in range(4, -1, -1):
        for i in range(nic):
            for j in range(njc):
                if colour5(i, j) != c:
                    continue
                acc = z[i, j].copy()
                for (di, dj) in SLOTS[1:]:
                    ni, nj = i + di, j + dj
                    if 0 <= ni < nic and 0 <= nj < njc and colour5(ni, nj) > c:
                        acc -= A[(i, j, ni, nj)] @ z[ni, nj]
                z[i, j] = Dinv[(i, j)] @ acc
    return z


def f16_scaled(a):
    """A block stored in float16 after scaling by its largest magnitude (the
    scale kept exactly), read back as float64."""
    s = np.abs(a).max()
    if s == 0.0:
        return a
    return (a / s).astype(np.float16).astype(np.float64) * s


def _ilu_ordered(B, idt, v, order, keep, transpose=False, f32=False, rnd=None):
    """Block ILU(0) of I/dt - J on the 13-point stencil with the cells in the
    given order, keeping only the couplings keep((i,j), (ii,jj)) accepts.
    f32: the factors (L, U and the inverted diagonal blocks) are stored
    rounded to float32 and later rows and the solves use the stored values,
    as the device factorisation does; a row's own L blocks enter its upper
    updates unrounded.  rnd: another storage rounding of a block (overrides
    f32)."""
    if rnd is None:
        rnd = (lambda a: a.astype(np.float32).astype(np.float64)) if f32 else (lambda a: a)
    v = np.asarray(v, dtype=np.float64)
    nic, njc, nv = v.shape
    pos = {c: n for n, c in enumerate(order)}
    rows = []
    for (i, j) in order:
        d = {}
        for s, (di, dj) in enumerate(SLOTS):
            ii, jj = i + di, j + dj
            if 0 <= ii < nic and 0 <= jj < njc and keep((i, j), (ii, jj)):
                blk = -B[i, j, s].copy()
                if s == 0:
                    blk += np.eye(nv) * idt[i, j]
                d[pos[(ii, jj)]] = blk
        rows.append(d)
    N = len(order)
    dinv = [None] * N
    for rr in range(N):
        d = rows[rr]
        for k in sorted(c for c in d if c < rr):
            d[k] = d[k] @ dinv[k]
            for jc, ukj in rows[k].items():
                if jc > k and jc in d:
                    d[jc] = d[jc] - d[k] @ ukj
        dinv[rr] = rnd(np.linalg.inv(d[rr]))
        for k in d:
            if k != rr:
                d[k] = rnd(d[k])
    vv = np.array([v[c] for c in order])
    if not transpose:
        y = np.zeros_like(vv)
        for rr in range(N):
            s = vv[rr].copy()
            for k, lk in rows[rr].items():
                if k < rr:
                    s -= lk @ y[k]
            y[rr] = s
        x = np.zeros_like(vv)
        for rr in reversed(range(N)):
            s = y[rr].copy()
            for k, uk in rows[rr].items():
                if k > rr:
                    s -= uk @ x[k]
            x[rr] = dinv[rr] @ s
    else:
        w = np.zeros_like(vv)
        for rr in range(N):
            s = vv[rr].copy()
            for k in rows[rr]:
                if k < rr:
                    s -= rows[k][rr].T @ w[k]
            w[rr] = dinv[rr].T @ s
        x = np.zeros_like(vv)
        for rr in reversed(range(N)):
            s = w[rr].copy()
            for k in rows[rr]:
                if k > rr:
                    s -= rows[k][rr].T @ x[k]
            x[rr] = s
    out = np.zeros_like(v)
    for n, c in enumerate(order):
        out[c] = x[n]
    return out


def ilu_natural(B, idt, v, transpose=False):
    """Block ILU(0) of I/dt - J in natural j-major order (o = j*nic + i)."""
    nic, njc = np.asarray(v).shape[:2]
    order = [(i, j) for j in range(njc) for i in range(nic)]
    return _ilu_ordered(B, idt, v, order, lambda a, b: True, transpose)


def ilu_chunked(B, idt, v, w, transpose=False, f32=True, nine_point=False, f16=False):
    """Chunked line block ILU(0): i-chunks of w columns, each ordered column
    by column with j fastest; couplings between chunks dropped; factors
    stored in float32 as on the device.  nine_point: the pattern is the 3x3
    box only (the +-2 slots, which hold only fill for a first-order
    Jacobian, are dropped).  f16: factors stored as float16 with a scale per
    block instead of float32."""
    nic, njc = np.asarray(v).shape[:2]
    order = [(i, j) for i in range(nic) for j in range(njc)]  # chunks are contiguous in i
    if nine_point:
        keep = lambda a, b: a[0] // w == b[0] // w and abs(a[0] - b[0]) <= 1 and abs(a[1] - b[1]) <= 1
    else:
        keep = lambda a, b: a[0] // w == b[0] // w
    return _ilu_ordered(B, idt, v, order, keep, transpose, f32, rnd=f16_scaled if f16 else None)
