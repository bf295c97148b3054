"""CPU restatement of NGA2's hypre_uns linsol (IJ/ParCSR path) and of the
algebraic multigrid the build runs behind it (numpy, float64).

Test infrastructure only: the tests, smoke() and the bench's CPU baseline leg
use it as the checker of libngapuns.so; the product never imports it.

The reference hands the matrix to HYPRE (absent from the reference tree and
from this image), so the solver algorithms below are restated from HYPRE's
published descriptions with the parameters hypre_uns sets, not from HYPRE's
source. The parts the reference itself computes are restated line by line:

- umap (hypre_uns_class.f90:562-607, prep_umap): a cell participates when
  |opr(1)| > 10 eps; participating cells are numbered 0.. in i-fastest order
  over the interior, ghosts take their periodic images (cfg.sync), all other
  entries are -1. The map is built once, at init, from opr as it is then.
- assemble (:186-230): one CSR row per participating cell in the same order;
  the columns of stencil entry st are ind(i+stc(st,:)); entries whose column
  is -1 are dropped; an entry repeating an earlier column (a periodic
  direction of one or two cells) is added into the earlier one. (The
  reference sets such a column to 0 and keeps it, which adds a stray entry in
  column 0; merging is the intended operation and gives the same matrix
  whenever every periodic direction has three or more cells.)
- solve (:430-500): rhs and the guess are read from the participating cells,
  the solution is written back to them only (other cells keep their values),
  then cfg.sync.

The multigrid hierarchy (BoomerAMG with hypre_uns's settings,
hypre_uns_class.f90:27-33 and :256-266):
- strength: classical, theta = 0.25; with a positive diagonal, j is a strong
  dependency of i when a_ij < theta * min_{k != i} a_ik < 0 (signs mirrored
  for a negative diagonal); a row whose |row sum| exceeds 0.9 |a_ii| has no
  strong dependencies (BoomerAMG's max row sum 0.9);
- PMIS coarsening (De Sterck, Yang, Heys 2006): measure = number of points
  strongly depending on i + a random number in [0, 1); points with measure
  < 1 are F; then repeatedly every undecided point whose measure beats all
  undecided strong neighbours (either direction) becomes C and every
  undecided point strongly depending on a new C point becomes F. The random
  numbers here are splitmix64 of the row index (HYPRE's own generator is
  not reproducible without HYPRE);
- extended+i interpolation (De Sterck, Falgout, Nolting, Yang 2008, eq. 4.9),
  truncated to the 4 largest weights per row with positive and negative
  weights rescaled separately to keep their sums (BoomerAMG's P max 4);
- Galerkin coarse operators A_c = P^T A P;
- coarsening stops at <= 512 rows (max_coarse), 25 levels, or when a level
  no longer shrinks; the coarsest level is solved by a dense inverse (with
  gamma 1 1^T added when its rows sum to zero, gamma = sum |a_ii| / n^2).
- smoother: l1-Jacobi, x += (f - A x) / d with d_i = sum_j |a_ij| (HYPRE's
  relax type 18; hypre_uns asks for hybrid Gauss-Seidel, type 6, which is
  sequential within a process and has no counterpart on a GPU), one sweep
  before and one after the coarse correction (NumSweeps 1), V-cycle from a
  zero guess (MaxIter 1, Tol 0 as preconditioner).

Krylov drivers, all stopping on the 2-norm relative residual ||r|| / ||b||
(TwoNorm 1, :251): PCG (pcg_amg; pcg with Jacobi scaling), right-
preconditioned restarted GMRES(5) (gmres_amg; gmres with Jacobi scaling,
gmres_kdim = 5, :28), right-preconditioned BiCGStab (bicgstab_amg), and the
multigrid as a stand-alone iteration (amg). pcg_parasail and gmres_pilut are
not provided.
"""
import numpy as np

EPS = np.finfo(np.float64).eps

AMG, PCG_AMG, GMRES_AMG, BICGSTAB_AMG, PCG, PCG_PARASAIL, GMRES, GMRES_PILUT = 1, 2, 3, 4, 5, 6, 7, 8

THETA = 0.25
MAX_ROW_SUM = 0.9
PMAX = 4
MAX_COARSE = 512
MAX_LEVELS = 25
KDIM = 5

C_PT, F_PT, U_PT = 1, -1, 0


# ------------------------------------------------------------------ mapping and matrix
def umap(cfg, opr):
    """ind (ghosted, int64) and the number of participating cells."""
    it = cfg.interior
    act = np.abs(opr[0][it]) > 10.0 * EPS
    flat = act.ravel(order="F")
    num = np.where(flat, np.cumsum(flat) - 1, -1).reshape(act.shape, order="F")
    ind = np.full(cfg.shape, -1, dtype=np.int64)
    ind[it] = num
    cfg.sync(ind)
    return ind, int(flat.sum())


def assemble(cfg, ind, opr, stc):
    """CSR (rowptr, col, val) of the participating rows, columns in stencil order."""
    no = cfg.no
    rowptr = [0]
    cols, vals = [], []
    nst = len(stc)
    for k in range(cfg.nz):
        for j in range(cfg.ny):
            for i in range(cfg.nx):
                I, J, K = i + no, j + no, k + no
                if ind[I, J, K] < 0:
                    continue
                rc, rv = [], []
                for st in range(nst):
                    c = int(ind[I + stc[st][0], J + stc[st][1], K + stc[st][2]])
                    if c < 0:
                        continue
                    v = float(opr[st, I, J, K])
                    if c in rc:
                        rv[rc.index(c)] += v
                    else:
                        rc.append(c)
                        rv.append(v)
                cols += rc
                vals += rv
                rowptr.append(len(cols))
    return np.array(rowptr, dtype=np.int64), np.array(cols, dtype=np.int64), np.array(vals, dtype=np.float64)


class CSR:
    def __init__(self, rowptr, col, val, ncol):
        self.rowptr = np.asarray(rowptr, dtype=np.int64)
        self.col = np.asarray(col, dtype=np.int64)
        self.val = np.asarray(val, dtype=np.float64)
        self.n = len(self.rowptr) - 1
        self.ncol = ncol

    def row(self, i):
        a, b = self.rowptr[i], self.rowptr[i + 1]
        return self.col[a:b], self.val[a:b]

    def matvec(self, x):
        y = np.zeros(self.n)
        for i in range(self.n):
            c, v = self.row(i)
            s = 0.0
            for q in range(len(c)):
                s += v[q] * x[c[q]]
            y[i] = s
        return y

    def dense(self):
        D = np.zeros((self.n, self.ncol))
        for i in range(self.n):
            c, v = self.row(i)
            for q in range(len(c)):
                D[i, c[q]] += v[q]
        return D

    def transpose(self):
        cnt = np.zeros(self.ncol + 1, dtype=np.int64)
        for c in self.col:
            cnt[c + 1] += 1
        rp = np.cumsum(cnt)
        pos = rp[:-1].copy()
        col = np.zeros(len(self.col), dtype=np.int64)
        val = np.zeros(len(self.col))
        for i in range(self.n):
            c, v = self.row(i)
            for q in range(len(c)):
                col[pos[c[q]]] = i
                val[pos[c[q]]] = v[q]
                pos[c[q]] += 1
        return CSR(rp, col, val, self.n)


def _diag(A, i):
    c, v = A.row(i)
    d = 0.0
    for q in range(len(c)):
        if c[q] == i:
            d += v[q]
    return d


# ------------------------------------------------------------------ setup
def strength(A, theta=THETA, max_row_sum=MAX_ROW_SUM):
    """Boolean per stored entry: i strongly depends on col (diagonal never)."""
    S = np.zeros(len(A.col), dtype=bool)
    for i in range(A.n):
        a, b = A.rowptr[i], A.rowptr[i + 1]
        c, v = A.col[a:b], A.val[a:b]
        diag = _diag(A, i)
        rsum = 0.0
        for x in v:
            rsum += float(x)
        if max_row_sum < 1.0 and abs(rsum) > abs(diag) * max_row_sum:
            continue
        off = c != i
        if not off.any():
            continue
        if diag >= 0.0:
            rs = v[off].min()
            if rs >= 0.0:
                continue
            thr = theta * rs
            for q in range(len(c)):
                if c[q] != i and v[q] < thr:
                    S[a + q] = True
        else:
            rs = v[off].max()
            if rs <= 0.0:
                continue
            thr = theta * rs
            for q in range(len(c)):
                if c[q] != i and v[q] > thr:
                    S[a + q] = True
    return S


_M64 = (1 << 64) - 1


def rand01(i):
    """splitmix64 of the row index -> [0, 1) with 53 bits."""
    z = (int(i) + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    z = z ^ (z >> 31)
    return (z >> 11) * (1.0 / 9007199254740992.0)


def strong_lists(A, S):
    dep = [[] for _ in range(A.n)]   # i strongly depends on j
    inf = [[] for _ in range(A.n)]   # j strongly depends on i (i influences j)
    for i in range(A.n):
        a, b = A.rowptr[i], A.rowptr[i + 1]
        for q in range(a, b):
            if S[q]:
                dep[i].append(int(A.col[q]))
                inf[int(A.col[q])].append(i)
    return dep, inf


def pmis(A, S):
    n = A.n
    dep, inf = strong_lists(A, S)
    w = np.array([len(inf[i]) + rand01(i) for i in range(n)])
    cf = np.zeros(n, dtype=np.int64)
    cf[w < 1.0] = F_PT
    nbr = [sorted(set(dep[i]) | set(inf[i])) for i in range(n)]
    while True:
        und = np.nonzero(cf == U_PT)[0]
        if len(und) == 0:
            break
        newc = [i for i in und if all(w[i] > w[j] for j in nbr[i] if cf[j] == U_PT)]
        for i in newc:
            cf[i] = C_PT
        newset = set(newc)
        for i in und:
            if cf[i] == U_PT and any(j in newset for j in dep[i]):
                cf[i] = F_PT
    return cf


def interp_ext_i(A, S, cf, pmax=PMAX):
    """P (n x nc) by extended+i interpolation, truncated to pmax per row."""
    n = A.n
    cmap = np.cumsum(cf == C_PT) - 1
    nc = int(np.sum(cf == C_PT))
    dep, _ = strong_lists(A, S)
    rows = []
    for i in range(n):
        if cf[i] == C_PT:
            rows.append(([int(cmap[i])], [1.0]))
            continue
        Cs = [j for j in dep[i] if cf[j] == C_PT]
        Fs = [j for j in dep[i] if cf[j] != C_PT]
        chat = list(Cs)
        for k in Fs:
            for j in dep[k]:
                if cf[j] == C_PT and j not in chat:
                    chat.append(j)
        if not chat:
            rows.append(([], []))
            continue
        cset = set(chat)
        sset = set(dep[i])
        w = {j: 0.0 for j in chat}
        diag = 0.0
        c, v = A.row(i)
        for q in range(len(c)):
            j = int(c[q])
            if j == i:
                diag += v[q]
            elif j in cset:
                w[j] += v[q]
            elif j not in sset:
                diag += v[q]  # weak, outside C-hat: lumped
        for q in range(len(c)):
            k = int(c[q])
            if k == i or k in cset or k not in sset:
                continue
            aik = v[q]
            ck, vk = A.row(k)
            sgn = 1.0 if _diag(A, k) >= 0.0 else -1.0
            den = 0.0
            for t in range(len(ck)):
                l = int(ck[t])
                if (l in cset or l == i) and vk[t] * sgn < 0.0:
                    den += vk[t]
            if den == 0.0:
                diag += aik
                continue
            dist = aik / den
            for t in range(len(ck)):
                l = int(ck[t])
                if vk[t] * sgn < 0.0:
                    if l in cset:
                        w[l] += dist * vk[t]
                    elif l == i:
                        diag += dist * vk[t]
        cols = [int(cmap[j]) for j in chat]
        vals = [-w[j] / diag for j in chat]
        cols, vals = truncate(cols, vals, pmax)
        order = np.argsort(cols, kind="stable")
        rows.append(([cols[o] for o in order], [vals[o] for o in order]))
    rowptr = [0]
    col, val = [], []
    for rc, rv in rows:
        col += rc
        val += rv
        rowptr.append(len(col))
    return CSR(rowptr, col, val, nc)


def truncate(cols, vals, pmax):
    if pmax <= 0 or len(cols) <= pmax:
        return cols, vals
    key = sorted(range(len(cols)), key=lambda q: (-abs(vals[q]), cols[q]))
    keep = sorted(key[:pmax])
    sp = sum(x for x in vals if x > 0)
    sn = sum(x for x in vals if x < 0)
    kp = sum(vals[q] for q in keep if vals[q] > 0)
    kn = sum(vals[q] for q in keep if vals[q] < 0)
    fp = sp / kp if kp != 0.0 else 1.0
    fn = sn / kn if kn != 0.0 else 1.0
    return [cols[q] for q in keep], [vals[q] * (fp if vals[q] > 0 else fn) for q in keep]


def spgemm(A, B):
    """C = A B, columns of each row ascending."""
    rowptr = [0]
    col, val = [], []
    for i in range(A.n):
        acc = {}
        c, v = A.row(i)
        for q in range(len(c)):
            cb, vb = B.row(c[q])
            for t in range(len(cb)):
                acc[int(cb[t])] = acc.get(int(cb[t]), 0.0) + v[q] * vb[t]
        ks = sorted(acc)
        col += ks
        val += [acc[k] for k in ks]
        rowptr.append(len(col))
    return CSR(rowptr, col, val, B.ncol)


def l1_norms(A):
    out = np.zeros(A.n)
    for i in range(A.n):
        s = 0.0
        for x in A.row(i)[1]:
            s += abs(float(x))
        out[i] = s
    return out


def coarse_inverse(A):
    M = A.dense()
    n = A.n
    rs = np.abs(M.sum(axis=1)).sum()
    dsum = np.abs(np.diag(M)).sum()
    singular = rs <= 1e-10 * dsum
    if singular:
        M = M + dsum / (n * n)
    return np.linalg.inv(M), singular


class Hierarchy:
    def __init__(self):
        self.A, self.P, self.R, self.cf, self.d = [], [], [], [], []
        self.Ainv = None
        self.singular = False


def build(A, theta=THETA, pmax=PMAX, max_coarse=MAX_COARSE, max_levels=MAX_LEVELS):
    H = Hierarchy()
    H.A.append(A)
    while A.n > max_coarse and len(H.A) < max_levels:
        S = strength(A, theta)
        cf = pmis(A, S)
        nc = int(np.sum(cf == C_PT))
        if nc == 0 or nc >= A.n:
            break
        P = interp_ext_i(A, S, cf, pmax)
        R = P.transpose()
        Ac = spgemm(R, spgemm(A, P))
        H.cf.append(cf)
        H.P.append(P)
        H.R.append(R)
        H.A.append(Ac)
        A = Ac
    H.d = [l1_norms(L) for L in H.A]
    H.Ainv, H.singular = coarse_inverse(H.A[-1])
    return H


# ------------------------------------------------------------------ cycle and Krylov
def vcycle(H, b, nu=1):
    """One V(nu, nu) cycle from a zero guess: x = M b."""
    L = len(H.A)
    f = [None] * L
    x = [None] * L
    f[0] = b
    for l in range(L - 1):
        A, d = H.A[l], H.d[l]
        x[l] = f[l] / d
        for _ in range(nu - 1):
            x[l] = x[l] + (f[l] - A.matvec(x[l])) / d
        r = f[l] - A.matvec(x[l])
        f[l + 1] = H.R[l].matvec(r)
    x[L - 1] = H.Ainv @ f[L - 1]
    for l in range(L - 2, -1, -1):
        A, d = H.A[l], H.d[l]
        x[l] = x[l] + H.P[l].matvec(x[l + 1])
        for _ in range(nu):
            x[l] = x[l] + (f[l] - A.matvec(x[l])) / d
    return x[0]


def _prec(method, H, A):
    if method in (AMG, PCG_AMG, GMRES_AMG, BICGSTAB_AMG):
        return lambda r: vcycle(H, r)
    dg = np.array([_diag(A, i) for i in range(A.n)])
    return lambda r: r / dg


def solve(method, A, H, b, x0, maxit, rcvg, kdim=KDIM):
    """(x, it, rerr): rerr the final true relative residual ||b - A x|| / ||b||."""
    bn = float(np.linalg.norm(b))
    x = np.array(x0, dtype=np.float64, copy=True)
    if bn == 0.0:
        return np.zeros_like(x), 0, 0.0
    M = _prec(method, H, A)
    res = lambda x: b - A.matvec(x)
    it = 0
    if method == AMG:
        r = res(x)
        while np.linalg.norm(r) / bn >= rcvg and it < maxit:
            x = x + M(r)
            r = res(x)
            it += 1
    elif method in (PCG, PCG_AMG):
        r = res(x)
        if np.linalg.norm(r) / bn >= rcvg:
            z = M(r)
            p = z.copy()
            rho = r @ z
            while it < maxit:
                q = A.matvec(p)
                alpha = rho / (p @ q)
                x = x + alpha * p
                r = r - alpha * q
                it += 1
                if np.linalg.norm(r) / bn < rcvg:
                    break
                z = M(r)
                rho1 = r @ z
                p = z + (rho1 / rho) * p
                rho = rho1
    elif method in (GMRES, GMRES_AMG):
        r = res(x)
        while np.linalg.norm(r) / bn >= rcvg and it < maxit:
            beta = np.linalg.norm(r)
            V = [r / beta]
            Z = []
            Hm = np.zeros((kdim + 1, kdim))
            m = 0
            for m in range(1, kdim + 1):
                z = M(V[-1])
                Z.append(z)
                w = A.matvec(z)
                for jj in range(m):
                    Hm[jj, m - 1] = V[jj] @ w
                    w = w - Hm[jj, m - 1] * V[jj]
                Hm[m, m - 1] = np.linalg.norm(w)
                it += 1
                e1 = np.zeros(m + 1)
                e1[0] = beta
                y = np.linalg.lstsq(Hm[:m + 1, :m], e1, rcond=None)[0]
                est = np.linalg.norm(e1 - Hm[:m + 1, :m] @ y)
                if Hm[m, m - 1] == 0.0 or est / bn < rcvg or it >= maxit:
                    break
                V.append(w / Hm[m, m - 1])
            for jj in range(len(y)):
                x = x + y[jj] * Z[jj]
            r = res(x)
    elif method == BICGSTAB_AMG:
        r = res(x)
        rh = r.copy()
        rho = alpha = omega = 1.0
        v = np.zeros_like(x)
        p = np.zeros_like(x)
        while np.linalg.norm(r) / bn >= rcvg and it < maxit:
            rho1 = rh @ r
            beta = (rho1 / rho) * (alpha / omega)
            p = r + beta * (p - omega * v)
            ph = M(p)
            v = A.matvec(ph)
            alpha = rho1 / (rh @ v)
            s = r - alpha * v
            it += 1
            if np.linalg.norm(s) / bn < rcvg:
                x = x + alpha * ph
                r = s
                break
            sh = M(s)
            t = A.matvec(sh)
            omega = (t @ s) / (t @ t)
            x = x + alpha * ph + omega * sh
            r = s - omega * t
            rho = rho1
    else:
        raise ValueError("method not provided")
    r = res(x)
    return x, it, float(np.linalg.norm(r) / bn)
