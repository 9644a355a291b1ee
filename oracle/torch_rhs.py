"""PyTorch FP64 restatement of compute_rhs_bssn at interior points, for
autograd.

Test infrastructure (like the rest of oracle/): the reference for the
hand-written adjoint.  torch.autograd differentiates this restatement, and
the vector-Jacobian products it gives are compared with the HIP adjoint
kernels.  Only tests and the benchmark's checking legs import it.

Scope: vacuum, Symmetry 0, points with +-3 room on every axis (there every
stencil is the full-order one: diff_new.f90:1084-1087, 1487-1507;
lopsidediff.f90:324-372; kodiss.f90:166-207), GAUGE 0 (bssn_rhs.f90:792-799)
and GAUGE 2 (:809-820).  The algebra follows bssn_rhs.f90:146-820 in tensor
form; the C restatement in bssn_oracle.c is its checked twin.

Arrays are indexed [k, j, i] (z, y, x), i.e. the reference's column-major
(i fastest) order read as C order.  Outputs cover [3, n-3) on every axis.
"""
from __future__ import annotations

import torch

SIDX = ((0, 1, 2), (1, 3, 4), (2, 4, 5))
PAIRS = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))
DIAG = (True, False, False, True, False, True)
WSYM = (1.0, 2.0, 2.0, 1.0, 2.0, 1.0)
CHI, TRK, DXX, AXX, GAMX, LAP, BETAX, DTSFX = 0, 1, 2, 8, 14, 17, 18, 21


class Stencils:
    """Interior stencils of one patch with spacing h = (hx, hy, hz)."""

    def __init__(self, h, eps):
        self.h = h
        self.d1 = [1.0 / 12.0 / h[a] for a in range(3)]
        self.d2 = [1.0 / 12.0 / (h[a] * h[a]) for a in range(3)]
        self.ko_c = [eps / 64.0 / h[a] for a in range(3)]
        self.dm = {(0, 1): 1.0 / 144.0 / (h[0] * h[1]), (0, 2): 1.0 / 144.0 / (h[0] * h[2]),
                   (1, 2): 1.0 / 144.0 / (h[1] * h[2])}

    @staticmethod
    def sh(f, d):
        """f shifted by d = (dx, dy, dz), restricted to the interior [3, n-3)."""
        nz, ny, nx = f.shape
        return f[3 + d[2]:nz - 3 + d[2], 3 + d[1]:ny - 3 + d[1], 3 + d[0]:nx - 3 + d[0]]

    @staticmethod
    def e(a, k):
        d = [0, 0, 0]
        d[a] = k
        return d

    def c(self, f):
        return self.sh(f, (0, 0, 0))

    def grad(self, f):
        s, e = self.sh, self.e
        return [self.d1[a] * ((s(f, e(a, -2)) - s(f, e(a, 2))) + 8.0 * (s(f, e(a, 1)) - s(f, e(a, -1))))
                for a in range(3)]

    def hess(self, f):
        s, e = self.sh, self.e
        out = []
        f0 = self.c(f)
        for (a, b) in PAIRS:
            if a == b:
                out.append(self.d2[a] * (16.0 * (s(f, e(a, -1)) + s(f, e(a, 1))) -
                                         (s(f, e(a, -2)) + s(f, e(a, 2))) - 30.0 * f0))
            else:
                def d(ka, kb):
                    v = [0, 0, 0]
                    v[a] += ka
                    v[b] += kb
                    return s(f, v)

                def row(kb):
                    return (d(-2, kb) - d(2, kb)) + 8.0 * (d(1, kb) - d(-1, kb))
                out.append(self.dm[(a, b)] * ((row(-2) - row(2)) + 8.0 * (row(1) - row(-1))))
        return out

    def adv(self, f, beta):
        s, e = self.sh, self.e
        f0 = 10.0 * self.c(f)
        r = 0.0
        for a in range(3):
            up = (s(f, e(a, 3)) - 3.0 * s(f, e(a, -1))) + (18.0 * s(f, e(a, 1)) - 6.0 * s(f, e(a, 2))) - f0
            dn = (s(f, e(a, -3)) - 3.0 * s(f, e(a, 1))) + (18.0 * s(f, e(a, -1)) - 6.0 * s(f, e(a, -2))) - f0
            r = r + self.d1[a] * (torch.clamp(beta[a], min=0.0) * up - torch.clamp(beta[a], max=0.0) * dn)
        return r

    def ko(self, f):
        s, e = self.sh, self.e
        c20 = 20.0 * self.c(f)
        r = 0.0
        for a in range(3):
            r = r + self.ko_c[a] * ((s(f, e(a, -3)) + s(f, e(a, 3))) - 6.0 * (s(f, e(a, -2)) + s(f, e(a, 2))) +
                                    15.0 * (s(f, e(a, -1)) + s(f, e(a, 1))) - c20)
        return r


def rhs(U, h, eps, gauge=0):
    """The 24 right-hand sides of compute_rhs_bssn at the interior points of
    one patch.  U: 24 tensors [nz, ny, nx] in the reference field order
    (chi, trK and the diagonal metric stored as offsets, as in the build)."""
    S = Stencils(h, eps)
    c = S.c
    chi1 = c(U[CHI]) + 1.0
    ichi = 1.0 / chi1
    K = c(U[TRK])
    alp = c(U[LAP]) + 1.0
    g = [c(U[DXX + s]) + (1.0 if DIAG[s] else 0.0) for s in range(6)]
    A = [c(U[AXX + s]) for s in range(6)]
    beta = [c(U[BETAX + a]) for a in range(3)]
    # inverse conformal metric (bssn_rhs.f90:194-201)
    cof = [g[3] * g[5] - g[4] * g[4], g[2] * g[4] - g[1] * g[5], g[1] * g[4] - g[2] * g[3],
           g[0] * g[5] - g[2] * g[2], g[1] * g[2] - g[0] * g[4], g[0] * g[3] - g[1] * g[1]]
    idet = 1.0 / (g[0] * cof[0] + g[1] * cof[1] + g[2] * cof[2])
    gu = [x * idet for x in cof]

    def G(T, i, j):
        return T[SIDX[i][j]]

    def tr(T):
        return gu[0] * T[0] + gu[3] * T[3] + gu[5] * T[5] + 2.0 * (gu[1] * T[1] + gu[2] * T[2] + gu[4] * T[4])

    AU = [[sum(G(gu, i, k) * G(A, k, j) for k in range(3)) for j in range(3)] for i in range(3)]
    Aup = [sum(AU[i][k] * G(gu, k, j) for k in range(3)) for (i, j) in PAIRS]
    out = [None] * 24

    # shift (:793-795; Lie and Laplacian terms of Gamma-tilde :260-353)
    db = []
    lapb, gdiv = [], [0.0, 0.0, 0.0]
    for a in range(3):
        f = U[BETAX + a]
        db.append(S.grad(f))
        hb = S.hess(f)
        lapb.append(tr(hb))
        for j in range(3):
            gdiv[j] = gdiv[j] + G(hb, j, a)
        out[BETAX + a] = 0.75 * c(U[DTSFX + a]) + S.adv(f, beta) + S.ko(f)
    Lb = [lapb[i] + (1.0 / 3.0) * sum(G(gu, i, j) * gdiv[j] for j in range(3)) for i in range(3)]
    divb = db[0][0] + db[1][1] + db[2][2]

    # metric (:162-191, :378-400)
    dg = [[None] * 6 for _ in range(3)]
    R = [None] * 6
    for s, (i, j) in enumerate(PAIRS):
        f = U[DXX + s]
        d = S.grad(f)
        for k in range(3):
            dg[k][s] = d[k]
        R[s] = -0.5 * tr(S.hess(f))
        lie = sum(G(g, i, k) * db[k][j] + G(g, j, k) * db[k][i] for k in range(3))
        out[DXX + s] = -2.0 * alp * A[s] + lie - (2.0 / 3.0) * g[s] * divb + S.adv(f, beta) + S.ko(f)
    G1 = [[0.5 * (dg[i][SIDX[j][k]] + dg[j][SIDX[i][k]] - dg[k][s]) for s, (i, j) in enumerate(PAIRS)]
          for k in range(3)]
    G2 = [[sum(G(gu, m, l) * G1[l][s] for l in range(3)) for s in range(6)] for m in range(3)]
    Gamxa = [tr(G2[m]) for m in range(3)]
    GA = [sum(WSYM[s] * G2[m][s] * Aup[s] for s in range(6)) for m in range(3)]
    # Ricci (:402-601)
    for i in range(3):
        Um = [[sum(G(gu, l, m) * G2[k][SIDX[l][i]] for l in range(3)) for k in range(3)] for m in range(3)]
        for j in range(3):
            P = sum(Um[m][k] * (G1[j][SIDX[k][m]] + G1[k][SIDX[m][j]]) for m in range(3) for k in range(3))
            H = sum(Um[m][k] * G1[j][SIDX[k][m]] for m in range(3) for k in range(3))
            gg = 0.5 * sum(Gamxa[l] * G1[i][SIDX[j][l]] for l in range(3))
            if j >= i:
                R[SIDX[i][j]] = R[SIDX[i][j]] + P + gg
            if j <= i:
                R[SIDX[j][i]] = R[SIDX[j][i]] + H + gg
    # Gamma-tilde derivatives into Ricci; advection + KO kept
    pGt = []
    for k in range(3):
        f = U[GAMX + k]
        d = S.grad(f)
        for s, (i, j) in enumerate(PAIRS):
            R[s] = R[s] + 0.5 * (G(g, k, i) * d[j] + G(g, k, j) * d[i])
        pGt.append(S.adv(f, beta) + S.ko(f))
    # chi (:160, :603-626)
    f = U[CHI]
    dchi = S.grad(f)
    hc = S.hess(f)
    out[CHI] = (2.0 / 3.0) * chi1 * (alp * K - divb) + S.adv(f, beta) + S.ko(f)
    fs = 0.0
    hh = [None] * 6
    for s, (i, j) in enumerate(PAIRS):
        hh[s] = hc[s] - (G2[0][s] * dchi[0] + G2[1][s] * dchi[1] + G2[2][s] * dchi[2])
        fs = fs + WSYM[s] * gu[s] * (hh[s] - 1.5 * ichi * dchi[i] * dchi[j])
    for s, (i, j) in enumerate(PAIRS):
        R[s] = R[s] + 0.5 * ichi * (hh[s] - 0.5 * ichi * dchi[i] * dchi[j] + g[s] * fs)
    # physical Christoffels (:632-650)
    u = [ichi * sum(G(gu, k, l) * dchi[l] for l in range(3)) for k in range(3)]
    P2 = [[None] * 6 for _ in range(3)]
    for k in range(3):
        for s, (i, j) in enumerate(PAIRS):
            d = 0.5 * g[s] * u[k]
            if k == i:
                d = d - 0.5 * ichi * dchi[j]
            if k == j:
                d = d - 0.5 * ichi * dchi[i]
            P2[k][s] = G2[k][s] + d
    # lapse (:629-664, :791)
    f = U[LAP]
    dalp = S.grad(f)
    ha = S.hess(f)
    out[LAP] = -2.0 * alp * K + S.adv(f, beta) + S.ko(f)
    DDa = [ha[s] - (P2[0][s] * dalp[0] + P2[1][s] * dalp[1] + P2[2][s] * dalp[2]) for s in range(6)]
    # trK derivatives
    f = U[TRK]
    dK = S.grad(f)
    pK = S.adv(f, beta) + S.ko(f)
    # Gamma-tilde rhs (:287-353) and the B driver (:792-820)
    gtr = []
    for i in range(3):
        Adchi = sum(G(Aup, i, j) * dchi[j] for j in range(3))
        Ada = sum(G(Aup, i, j) * dalp[j] for j in range(3))
        src = sum(G(gu, i, j) * ((2.0 / 3.0) * dK[j]) for j in range(3))
        ab = sum(Gamxa[j] * db[i][j] for j in range(3))
        gtr.append(-2.0 * Ada + 2.0 * alp * (GA[i] - 1.5 * ichi * Adchi - src) + Lb[i] +
                   (2.0 / 3.0) * Gamxa[i] * divb - ab)
        out[GAMX + i] = gtr[i] + pGt[i]
    if gauge == 2:
        q = sum(G(gu, i, j) * dchi[i] * dchi[j] for i in range(3) for j in range(3))
        w = 1.0 - torch.sqrt(chi1)
        eta = 0.655 * torch.sqrt(q * ichi) / (w * w)
    else:
        eta = 2.0
    for a in range(3):
        f = U[DTSFX + a]
        out[DTSFX + a] = gtr[a] - eta * c(f) + S.adv(f, beta) + S.ko(f)
    # trK and A_ij (:666-787)
    trAA = sum(AU[i][j] * AU[j][i] for i in range(3) for j in range(3))
    trDDa = tr(DDa)
    out[TRK] = -chi1 * trDDa + alp * ((1.0 / 3.0) * K * K + trAA) + pK
    f0 = (2.0 / 3.0) * K * K - trAA
    F = -(1.0 / 3.0) * (trDDa + alp * ichi * f0)
    for s, (i, j) in enumerate(PAIRS):
        lie = sum(G(A, i, k) * db[k][j] + G(A, j, k) * db[k][i] for k in range(3))
        AA = sum(G(A, i, k) * AU[k][j] for k in range(3))
        k_ = chi1 * (alp * R[s] - DDa[s] - g[s] * F) + alp * (K * A[s] - 2.0 * AA) + lie - (2.0 / 3.0) * A[s] * divb
        f = U[AXX + s]
        out[AXX + s] = k_ + S.adv(f, beta) + S.ko(f)
    return out


def enforce(U):
    """enforce_ga (enforce_algebra.f90:72-127) on every point: unit
    determinant of the conformal metric, trace-free A (trace taken with the
    cofactors of the rescaled metric)."""
    g = [U[DXX + s] + (1.0 if DIAG[s] else 0.0) for s in range(6)]
    det = (g[0] * g[3] * g[5] + g[1] * g[4] * g[2] + g[2] * g[1] * g[4] - g[2] * g[3] * g[2] -
           g[1] * g[1] * g[5] - g[0] * g[4] * g[4])
    sc = 1.0 / torch.pow(det, 1.0 / 3.0)
    p = [x * sc for x in g]
    cof = [p[3] * p[5] - p[4] * p[4], p[2] * p[4] - p[1] * p[5], p[1] * p[4] - p[3] * p[2],
           p[0] * p[5] - p[2] * p[2], p[1] * p[2] - p[0] * p[4], p[0] * p[3] - p[1] * p[1]]
    A = [U[AXX + s] for s in range(6)]
    trA = sum(WSYM[s] * cof[s] * A[s] for s in range(6))
    out = list(U)
    for s in range(6):
        out[DXX + s] = p[s] - (1.0 if DIAG[s] else 0.0)
        out[AXX + s] = A[s] - (1.0 / 3.0) * p[s] * trA
    return out


def step(U, dt, h, eps, chitiny=1e-5, gauge=0):
    """One RK4 step of bssn_class::Step for one block with frozen ghost
    zones of width 3 (bssn_class.C:3042-3163, 3401-3526): enforce_ga on the
    state, four right-hand sides, rungekutta4_rout's combinations
    (rungekutta4_rout.f90:110-128), lowerboundset on chi (fmisc.f90:1644-1657)
    and the next call's enforce_ga on each stage output; ghost points take
    the state's values (with the same bound and enforce)."""
    tinny = chitiny - 1.0
    I = (slice(3, -3), slice(3, -3), slice(3, -3))
    Y0 = enforce(U)
    c = (0.5 * dt, 0.5 * dt, dt, dt / 6.0)
    Y = Y0
    acc = None
    for st in range(4):
        k = rhs(Y, h, eps, gauge)
        if st == 0:
            acc = list(k)
        elif st < 3:
            acc = [a + 2.0 * x for a, x in zip(acc, k)]
        z = [Y0[v][I] + c[st] * (k[v] + acc[v] if st == 3 else k[v]) for v in range(24)]
        z[CHI] = torch.where(z[CHI] < tinny, torch.full_like(z[CHI], tinny), z[CHI])
        ghost = list(Y0)
        ghost[CHI] = torch.where(Y0[CHI] < tinny, torch.full_like(Y0[CHI], tinny), Y0[CHI])
        if st < 3:
            z = enforce(z)
            ghost = enforce(ghost)
        Yn = []
        for v in range(24):
            t = ghost[v].clone()
            t[I] = z[v]
            Yn.append(t)
        Y = Yn
    return Y


def interior_vjp(U_np, lam_np, h, eps, gauge=0):
    """d/dU of sum_v sum_p lam_v(p) rhs_v(p) over the interior points, by
    torch autograd.  U_np: 24 arrays [nz, ny, nx]; lam_np: 24 arrays of the
    interior shape.  Returns (rhs list, vjp list) as numpy arrays."""
    U = [torch.tensor(u, dtype=torch.float64, requires_grad=True) for u in U_np]
    out = rhs(U, h, eps, gauge)
    lam = [torch.tensor(x, dtype=torch.float64) for x in lam_np]
    grads = torch.autograd.grad(out, U, grad_outputs=lam, allow_unused=True)
    gr = [g.numpy() if g is not None else torch.zeros_like(u).numpy() for g, u in zip(grads, U)]
    return [o.detach().numpy() for o in out], gr


__all__ = ["rhs", "enforce", "step", "interior_vjp", "Stencils"]
