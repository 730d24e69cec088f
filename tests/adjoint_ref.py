"""Independent numpy restatement of the axisymmetric explicit solve and of its
adjoint w.r.t. (A, B, n) -- test infrastructure for the GPU adjoint.

The forward is the smooth idealisation of the reference's increment
(FE/ExplicitSolverCH.C:115-278, FE/Element.C:164-403): closed-form 2x2 polar
decomposition instead of the Jacobi iteration, radial return converged to
round-off instead of the 1e-8 stop.  With thermal=dict(m, T0, Tm, cp, tq) the
yield stress carries the Johnson-Cook temperature factor of
FE/IsoHardJohnsonCook.h:51-68, sigma_y = (A + B evp^n)(1 - T*^m), with
T* = (T - T0)/(Tm - T0) at the point's temperature before the increment, and
the point heats adiabatically by tq sigma_y d(evp) / (rho cp) after the
return (no rate term).  Its adjoint is written out step by step
(nodal update, assembly, element kinematics, polar decomposition, return
mapping, time-step control) and is checked against central differences of this
forward, which pins the derivation independently of the GPU code.
"""
import numpy as np

R3 = np.sqrt(3.0)
S23, S32 = np.sqrt(2.0 / 3.0), np.sqrt(1.5)


def sg0(k): q = k & 3; return 1 if q in (1, 2) else -1
def sg1(k): return 1 if (k & 3) >= 2 else -1


c0, c1 = (3 - R3) / 12, (3 + R3) / 12
n0, n1, n2 = (2 - R3) / 6, 1 / 6, (2 + R3) / 6
G = np.zeros((4, 4, 2))  # [pt][k][dir]
N = np.zeros((4, 4))
for pt in range(4):
    for k in range(4):
        m0 = c1 if sg1(pt) == sg1(k) else c0
        m1 = c1 if sg0(pt) == sg0(k) else c0
        G[pt, k, 0] = m0 if sg0(k) > 0 else -m0
        G[pt, k, 1] = m1 if sg1(k) > 0 else -m1
        ag = (sg0(pt) == sg0(k)) + (sg1(pt) == sg1(k))
        N[pt, k] = n2 if ag == 2 else (n1 if ag == 1 else n0)


class Model:
    def __init__(self, mesh, E=115e9, nu=0.3, rho=8930.0, rho_b=1.0, safety=0.9, freq=10, thermal=None):
        self.x0 = mesh["x"][:, :2].copy()
        self.conn = mesh["conn"].astype(np.int64)
        self.bc = mesh["bc"]
        self.v0 = mesh["v0"][:, :2].copy()
        self.K = E / (3 * (1 - 2 * nu))
        self.Gm = E / (2 * (1 + nu))
        rb = rho_b
        self.am = (2 * rb - 1) / (1 + rb)
        self.beta = (5 - 3 * rb) / ((1 + rb) ** 2 * (2 - rb))
        self.gam = 1.5 - self.am
        self.om = np.sqrt((12 * (1 + rb) ** 3 * (2 - rb)) / (10 + 15 * rb - rb * rb + rb ** 3 - rb ** 4))
        self.safety, self.freq = safety, freq
        self.cs = np.sqrt((E * (1 - nu)) / (rho * (1 + nu) * (1 - 2 * nu)))
        self.rho = rho
        # restrained masks per comp
        self.fix = np.stack([(self.bc & 1) > 0, (self.bc & 2) > 0], 1)
        self.mass = self.lumped_mass(self.x0)
        self.thermal = thermal
        if thermal:
            self.T0, self.Tm, self.mexp = thermal["T0"], thermal["Tm"], thermal["m"]
            self.kheat = thermal["tq"] / (rho * thermal["cp"])

    def theta(self, T):
        """temperature factor and its derivative"""
        if not self.thermal:
            return np.ones_like(T), np.zeros_like(T)
        ts = (T - self.T0) / (self.Tm - self.T0)
        inside = (ts > 0) & (ts < 1)
        tsc = np.where(inside, ts, 0.5)
        th = np.where(ts <= 0, 1.0, np.where(ts >= 1, 0.0, 1 - tsc ** self.mexp))
        dth = np.where(inside, -self.mexp * tsc ** (self.mexp - 1) / (self.Tm - self.T0), 0.0)
        return th, dth

    def geom(self, x):
        X = x[self.conn]  # (ne,4,2)
        J = np.einsum('pkd,ekc->epcd', G, X)  # J[c][d] = sum_k G[p,k,d] x_k[c]
        det = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
        inv = np.empty_like(J)
        inv[..., 0, 0] = J[..., 1, 1] / det
        inv[..., 0, 1] = -J[..., 0, 1] / det
        inv[..., 1, 0] = -J[..., 1, 0] / det
        inv[..., 1, 1] = J[..., 0, 0] / det
        Rr = np.einsum('pk,ek->ep', N, X[..., 0])
        d = np.einsum('pki,epij->epkj', G, inv)  # dN_k/dx_j
        return J, det, inv, Rr, d

    def lumped_mass(self, x):
        J, det, inv, Rr, d = self.geom(x)
        W = det * 2 * np.pi * Rr
        me = self.rho * np.einsum('pk,ep->ek', N, W)
        m = np.zeros(len(x))
        np.add.at(m, self.conn, me)
        return m

    def courant(self, x):
        X = x[self.conn]
        x31 = X[:, 2, 0] - X[:, 0, 0]; x24 = X[:, 1, 0] - X[:, 3, 0]
        y42 = X[:, 3, 1] - X[:, 1, 1]; y31 = X[:, 2, 1] - X[:, 0, 1]
        L = (x31 * y42 + x24 * y31) / np.sqrt(2 * (x24 ** 2 + y42 ** 2 + x31 ** 2 + y31 ** 2))
        return L / self.cs

    def courant_grad(self, x, e):
        """d(L_e/cs)/dx of element e's 4 nodes: (4,2)"""
        X = x[self.conn[e]]
        x31 = X[2, 0] - X[0, 0]; x24 = X[1, 0] - X[3, 0]
        y42 = X[3, 1] - X[1, 1]; y31 = X[2, 1] - X[0, 1]
        num = x31 * y42 + x24 * y31
        q = 2 * (x24 ** 2 + y42 ** 2 + x31 ** 2 + y31 ** 2)
        den = np.sqrt(q)
        # L = num/den ; dL/dv = dnum/den - num/(2 q den) dq
        dn = dict(x31=y42, y42=x31, x24=y31, y31=x24)
        dq = dict(x31=4 * x31, y42=4 * y42, x24=4 * x24, y31=4 * y31)
        dv = {k: dn[k] / den - num / (2 * q * den) * dq[k] for k in dn}
        g = np.zeros((4, 2))
        g[2, 0] += dv['x31']; g[0, 0] -= dv['x31']
        g[1, 0] += dv['x24']; g[3, 0] -= dv['x24']
        g[3, 1] += dv['y42']; g[1, 1] -= dv['y42']
        g[2, 1] += dv['y31']; g[0, 1] -= dv['y31']
        return g / self.cs

    # ---------------- element ----------------
    def elem_forward(self, x, du, S, evp, p, th, T=None, keep=False):
        A, B, n = th
        if T is None:
            T = np.zeros_like(evp)
        tf, dtf = self.theta(T)
        J, det, inv, Rr, d = self.geom(x)
        U = du[self.conn]  # (ne,4,2)
        ne = len(self.conn)
        F = np.zeros((ne, 4, 3, 3))
        F[..., 0, 0] = 1 + np.einsum('epk,ek->ep', d[..., 0], U[..., 0])
        F[..., 0, 1] = np.einsum('epk,ek->ep', d[..., 1], U[..., 0])
        F[..., 1, 0] = np.einsum('epk,ek->ep', d[..., 0], U[..., 1])
        F[..., 1, 1] = 1 + np.einsum('epk,ek->ep', d[..., 1], U[..., 1])
        vr = np.einsum('pk,ek->ep', N, U[..., 0])
        F[..., 2, 2] = 1 + vr / Rr
        # polar, 2x2 closed form
        a_ = F[..., 0, 0] + F[..., 1, 1]; b_ = F[..., 1, 0] - F[..., 0, 1]
        rr = np.hypot(a_, b_)
        c, s = a_ / rr, b_ / rr
        Rm = np.zeros_like(F)
        Rm[..., 0, 0] = c; Rm[..., 0, 1] = -s; Rm[..., 1, 0] = s; Rm[..., 1, 1] = c; Rm[..., 2, 2] = 1
        U2 = np.einsum('...ji,...jk->...ik', Rm[..., :2, :2], F[..., :2, :2])
        U2 = 0.5 * (U2 + np.swapaxes(U2, -1, -2))
        lam, Q = np.linalg.eigh(U2)
        lnU2 = np.einsum('...ij,...j,...kj->...ik', Q, np.log(lam), Q)
        dE = np.zeros_like(F)
        dE[..., :2, :2] = lnU2
        dE[..., 2, 2] = np.log(F[..., 2, 2])
        # stress update
        tr = np.trace(dE, axis1=-2, axis2=-1)
        dev = dE - tr[..., None, None] / 3 * np.eye(3)
        St = S + 2 * self.Gm * dev
        s_ = np.sqrt(np.sum(St ** 2, axis=(-1, -2)))
        Strial = S32 * s_
        sig0 = (A + B * evp ** n) * tf
        plastic = Strial > sig0
        gam = np.zeros_like(evp)
        if plastic.any():
            g = np.zeros(plastic.sum())
            ev0 = evp[plastic]; St_ = Strial[plastic]; tp = tf[plastic]
            for it in range(100):
                ev = ev0 + S23 * g
                phi = S32 * 2 * self.Gm * g + (A + B * ev ** n) * tp - St_
                dphi = S32 * 2 * self.Gm + S23 * n * B * ev ** (n - 1) * tp
                dg = -phi / dphi
                dg = np.where((dg < 0) & (-dg > g), -0.9 * g, dg)
                g = g + dg
                if np.all(np.abs(phi) <= 1e-15 * (A + B * ev ** n) * tp):
                    break
            gam[plastic] = g
        cfac = np.where(plastic, 1 - 2 * self.Gm * gam / np.where(s_ > 0, s_, 1), 1.0)
        Sp = cfac[..., None, None] * St
        evp_new = evp + S23 * gam
        T_new = T.copy()
        if self.thermal:
            T_new = np.where(plastic, T + self.kheat * (A + B * evp_new ** n) * tf * S23 * gam, T)
        pres = np.mean(tr, axis=1)
        p_new = p + self.K * pres
        Srot = np.einsum('...ij,...jk,...lk->...il', Rm, Sp, Rm)
        Sig = Srot + p_new[:, None, None, None] * np.eye(3)
        W = det * 2 * np.pi * Rr
        fe = np.zeros((ne, 4, 2))
        # f_L0 = -sum_q (d_qL0 Sig00 + d_qL1 Sig10 + N_qL Sig22/Rr_q) W_q
        fe[..., 0] = -np.einsum('eqL,eq->eL', d[..., 0], Sig[..., 0, 0] * W) \
            - np.einsum('eqL,eq->eL', d[..., 1], Sig[..., 1, 0] * W) \
            - np.einsum('qL,eq->eL', N, Sig[..., 2, 2] * W / Rr)
        fe[..., 1] = -np.einsum('eqL,eq->eL', d[..., 0], Sig[..., 0, 1] * W) \
            - np.einsum('eqL,eq->eL', d[..., 1], Sig[..., 1, 1] * W)
        out = (Srot, evp_new, p_new, fe, T_new)
        if keep:
            tape = dict(J=J, det=det, inv=inv, Rr=Rr, d=d, U=U, F=F, vr=vr, a_=a_, b_=b_, rr=rr, c=c, s=s, Rm=Rm,
                        U2=U2, lam=lam, Q=Q, dE=dE, St=St, s_=s_, Strial=Strial, plastic=plastic, gam=gam,
                        cfac=cfac, Sp=Sp, Sig=Sig, W=W, evp=evp, tf=tf, dtf=dtf)
            return out, tape
        return out

    def elem_backward(self, x, du, S, evp, p, th, fbar, Sbar, evbar, pbar, Temp=None, Tbar=None):
        """VJP of elem_forward: returns xbar (nn,2), dubar (nn,2), Sbar_in, evbar_in, pbar_in, thbar (3),
        Tbar_in."""
        A, B, n = th
        if Temp is None:
            Temp = np.zeros_like(evp)
        if Tbar is None:
            Tbar = np.zeros_like(evp)
        (Srot, evp_new, p_new, fe, _), T = self.elem_forward(x, du, S, evp, p, th, Temp, keep=True)
        d, W, Rr, Sig, Rm, Sp = T['d'], T['W'], T['Rr'], T['Sig'], T['Rm'], T['Sp']
        ne = len(self.conn)
        Sigb = np.zeros((ne, 4, 3, 3))
        fb0, fb1 = fbar[..., 0], fbar[..., 1]  # (ne,4 nodes)
        Sigb[..., 0, 0] = -np.einsum('eL,eqL->eq', fb0, d[..., 0]) * W
        Sigb[..., 1, 0] = -np.einsum('eL,eqL->eq', fb0, d[..., 1]) * W
        Sigb[..., 2, 2] = -np.einsum('eL,qL->eq', fb0, N) * W / Rr
        Sigb[..., 0, 1] = -np.einsum('eL,eqL->eq', fb1, d[..., 0]) * W
        Sigb[..., 1, 1] = -np.einsum('eL,eqL->eq', fb1, d[..., 1]) * W
        db = np.zeros_like(d)  # (e,q,L,j)
        db[..., 0] = -(fb0[:, None, :] * Sig[..., 0, 0][..., None] + fb1[:, None, :] * Sig[..., 0, 1][..., None]) * W[..., None]
        db[..., 1] = -(fb0[:, None, :] * Sig[..., 1, 0][..., None] + fb1[:, None, :] * Sig[..., 1, 1][..., None]) * W[..., None]
        Wb = -(np.einsum('eL,eqL->eq', fb0, d[..., 0]) * Sig[..., 0, 0] + np.einsum('eL,eqL->eq', fb0, d[..., 1]) * Sig[..., 1, 0]
               + np.einsum('eL,qL->eq', fb0, N) * Sig[..., 2, 2] / Rr
               + np.einsum('eL,eqL->eq', fb1, d[..., 0]) * Sig[..., 0, 1] + np.einsum('eL,eqL->eq', fb1, d[..., 1]) * Sig[..., 1, 1])
        Rrb = np.einsum('eL,qL->eq', fb0, N) * Sig[..., 2, 2] * W / Rr ** 2
        detb = Wb * 2 * np.pi * Rr
        Rrb += Wb * 2 * np.pi * T['det']
        # Sig = Srot + p I
        Srotb = Sbar + Sigb
        pb_tot = pbar + np.einsum('eqii->e', Sigb)
        # Srot = R Sp R^T
        Spb = np.einsum('...ji,...jk,...kl->...il', Rm, Srotb, Rm)
        Rb = np.einsum('...ij,...jk,...lk->...il', Srotb, Rm, Sp) + np.einsum('...ji,...jk,...kl->...il', Srotb, Rm, Sp)
        # p_new = p + K mean(tr)
        pb_in = pb_tot.copy()
        trb = np.repeat((self.K / 4 * pb_tot)[:, None], 4, axis=1)
        # radial return backward
        St, s_, gam, plastic, cfac = T['St'], T['s_'], T['gam'], T['plastic'], T['cfac']
        evb_in = evbar.copy()
        Stb = cfac[..., None, None] * Spb
        cb = np.sum(Spb * St, axis=(-1, -2))
        tf, dtf = T['tf'], T['dtf']
        Tb_in = Tbar.copy()
        gb = np.where(plastic, -2 * self.Gm / np.where(s_ > 0, s_, 1) * cb + S23 * evbar, 0.0)
        sb = np.where(plastic, 2 * self.Gm * gam / np.where(s_ > 0, s_, 1) ** 2 * cb, 0.0)
        ev = evp + S23 * gam
        with np.errstate(all='ignore'):
            evn = ev ** n
            base = A + B * evn
            hb = n * B * ev ** (n - 1)
            lev = np.log(ev)
            thA = np.zeros_like(evp); thB = np.zeros_like(evp); thN = np.zeros_like(evp)
            if self.thermal:
                # T' = T + k base(ev) theta(T) sqrt(2/3) gamma on the plastic branch
                kq = np.where(plastic, Tbar * self.kheat, 0.0)
                dep = S23 * gam
                gb = gb + np.where(plastic, kq * tf * S23 * (hb * dep + base), 0.0)
                evb_in += np.where(plastic, kq * tf * hb * dep, 0.0)
                Tb_in += np.where(plastic, kq * base * dtf * dep, 0.0)
                thA += np.where(plastic, kq * tf * dep, 0.0)
                thB += np.where(plastic, kq * tf * dep * evn, 0.0)
                thN += np.where(plastic, kq * tf * dep * B * evn * lev, 0.0)
            dphi = S32 * 2 * self.Gm + S23 * hb * tf
            w = np.where(plastic, gb / dphi, 0.0)
            Strialb = w
            evb_in += np.where(plastic, -w * hb * tf, 0.0)
            Tb_in += np.where(plastic, -w * base * dtf, 0.0)
            thA += np.where(plastic, -w * tf, 0.0)
            thB += np.where(plastic, -w * tf * evn, 0.0)
            thN += np.where(plastic, -w * tf * B * evn * lev, 0.0)
            thb = np.array([np.sum(thA), np.sum(thB), np.sum(thN)])
        sb = sb + S32 * Strialb
        Stb = Stb + np.where(plastic, sb / np.where(s_ > 0, s_, 1), 0.0)[..., None, None] * St
        Sb_in = Stb.copy()
        X = 2 * self.Gm * Stb
        dEb = X - np.trace(X, axis1=-2, axis2=-1)[..., None, None] / 3 * np.eye(3)
        dEb = dEb + trb[..., None, None] * np.eye(3)
        dEb = 0.5 * (dEb + np.swapaxes(dEb, -1, -2))
        # polar backward
        F = T['F']; Q = T['Q']; lam = T['lam']
        Fb = np.zeros_like(F)
        Fb[..., 2, 2] = dEb[..., 2, 2] / F[..., 2, 2]
        Lb = dEb[..., :2, :2]
        M = np.einsum('...ji,...jk,...kl->...il', Q, Lb, Q)
        l1, l2 = lam[..., 0], lam[..., 1]
        with np.errstate(all='ignore'):
            dd = np.where(np.abs(l1 - l2) > 1e-12 * l1, (np.log(l1) - np.log(l2)) / (l1 - l2), 1 / np.sqrt(l1 * l2))
        Fdd = np.zeros_like(M)
        Fdd[..., 0, 0] = 1 / l1; Fdd[..., 1, 1] = 1 / l2; Fdd[..., 0, 1] = dd; Fdd[..., 1, 0] = dd
        U2b = np.einsum('...ij,...jk,...lk->...il', Q, M * Fdd, Q)
        R2 = Rm[..., :2, :2]
        F2 = F[..., :2, :2]
        Fb[..., :2, :2] += np.einsum('...ij,...jk->...ik', R2, U2b)
        R2b = Rb[..., :2, :2] + np.einsum('...ij,...kj->...ik', F2, U2b)
        c, s = T['c'], T['s']
        phib = R2b[..., 0, 0] * (-s) + R2b[..., 0, 1] * (-c) + R2b[..., 1, 0] * c + R2b[..., 1, 1] * (-s)
        a_, b_, rr = T['a_'], T['b_'], T['rr']
        Fb[..., 0, 0] += -phib * b_ / rr ** 2
        Fb[..., 1, 1] += -phib * b_ / rr ** 2
        Fb[..., 1, 0] += phib * a_ / rr ** 2
        Fb[..., 0, 1] += -phib * a_ / rr ** 2
        # F backward
        Uel = T['U']
        db[..., 0] += Fb[..., 0, 0][..., None] * Uel[:, None, :, 0] + Fb[..., 1, 0][..., None] * Uel[:, None, :, 1]
        db[..., 1] += Fb[..., 0, 1][..., None] * Uel[:, None, :, 0] + Fb[..., 1, 1][..., None] * Uel[:, None, :, 1]
        Ub = np.zeros_like(Uel)
        Ub[..., 0] = np.einsum('eq,eqk->ek', Fb[..., 0, 0], d[..., 0]) + np.einsum('eq,eqk->ek', Fb[..., 0, 1], d[..., 1]) \
            + np.einsum('eq,qk->ek', Fb[..., 2, 2] / Rr, N)
        Ub[..., 1] = np.einsum('eq,eqk->ek', Fb[..., 1, 0], d[..., 0]) + np.einsum('eq,eqk->ek', Fb[..., 1, 1], d[..., 1])
        Rrb += -Fb[..., 2, 2] * T['vr'] / Rr ** 2
        # d = G inv
        invb = np.einsum('pki,epkj->epij', G, db)
        inv = T['inv']
        Jb = -np.einsum('...ji,...jk,...lk->...il', inv, invb, inv)
        Jm = T['J']
        Jb[..., 0, 0] += detb * Jm[..., 1, 1]; Jb[..., 1, 1] += detb * Jm[..., 0, 0]
        Jb[..., 0, 1] -= detb * Jm[..., 1, 0]; Jb[..., 1, 0] -= detb * Jm[..., 0, 1]
        # J[c][d] = sum_k G[p,k,d] x_k[c]
        Xb = np.einsum('epcd,pkd->ekc', Jb, G)
        Xb[..., 0] += np.einsum('ep,pk->ek', Rrb, N)
        xb = np.zeros((len(x), 2)); dub = np.zeros((len(x), 2))
        np.add.at(xb, self.conn, Xb)
        np.add.at(dub, self.conn, Ub)
        return xb, dub, Sb_in, evb_in, pb_in, thb, Tb_in

    # ---------------- nodal ----------------
    def nodal(self, x, v, a, f, dt):
        s1, s2, s3, s4 = dt * (0.5 - self.beta), dt * (1 - self.gam), dt * self.gam, dt * dt * self.beta
        dun = (a * s1 + v) * dt
        vn = a * s2 + v
        aa = (f / self.mass[:, None] - self.am * a) / (1 - self.am)
        vn = vn + s3 * aa
        dun = dun + s4 * aa
        dun[self.fix] = 0; vn[self.fix] = 0; aa[self.fix] = 0
        return x + dun, vn, aa, dun

    def nodal_backward(self, v, a, aa, dt, Xb, Ub, Vb, Ab):
        s1, s2, s3, s4 = dt * (0.5 - self.beta), dt * (1 - self.gam), dt * self.gam, dt * dt * self.beta
        free = ~self.fix
        dunb = (Xb + Ub) * free
        Vb = Vb * free; Ab = Ab * free
        aab = Ab + Vb * s3 + dunb * s4
        s3b = np.sum(Vb * aa); s4b = np.sum(dunb * aa)
        an0b = aab / (1 - self.am)
        ab = -self.am / (1 - self.am) * aab + Vb * s2 + dunb * s1 * dt
        vb = Vb + dunb * dt
        s2b = np.sum(Vb * a); s1b = np.sum(dunb * a * dt)
        dtb = np.sum(dunb * (a * s1 + v))
        dtb += s1b * (0.5 - self.beta) + s2b * (1 - self.gam) + s3b * self.gam + s4b * 2 * dt * self.beta
        fb = an0b / self.mass[:, None]
        return Xb.copy(), vb, ab, fb, dtb

    # ---------------- run ----------------
    def run(self, th, nsteps, tape=False):
        x = self.x0.copy(); v = self.v0.copy(); a = np.zeros_like(x); du = np.zeros_like(x)
        ne = len(self.conn)
        S = np.zeros((ne, 4, 3, 3)); evp = np.full((ne, 4), 1e-10); p = np.zeros(ne)
        Temp = np.full((ne, 4), self.T0 if self.thermal else 0.0)
        t = 0.0
        crit = self.courant(x)
        dt = self.safety * self.om / (2 / crit.min())
        ts, lts, Cc, Ls = 0.0, 0.0, dt, 0.0
        hist = []
        for k in range(nsteps):
            sc = dict(t=t, dt_in=dt, ts=ts, lts=lts, C=Cc, L=Ls, rec=False, amin=-1)
            if k > 0:
                if k % self.freq == 0:
                    cr = self.courant(x)
                    e = int(np.argmin(cr))
                    dt = self.safety * self.om / (2 / cr[e])
                    lts, Ls, ts, Cc = ts, Cc, t, dt
                    sc.update(rec=True, amin=e)
                elif ts != lts:
                    dx = (t - lts) / (ts - lts)
                    dt = dx * Cc + (1 - dx) * Ls
            state = (x, v, a, du, S, evp, p, Temp)
            S, evp, p, fe, Temp = self.elem_forward(x, du, S, evp, p, th, Temp)
            f = np.zeros_like(x)
            np.add.at(f, self.conn, fe)
            x, v, a, du = self.nodal(x, v, a, f, dt)
            t = t + dt
            sc['dt'] = dt
            if tape:
                hist.append((state, sc, a.copy()))
        return x, hist

    def grad(self, th, nsteps, xt):
        xN, hist = self.run(th, nsteps, tape=True)
        Xb = 2 * (xN - xt); Ub = np.zeros_like(Xb); Vb = np.zeros_like(Xb); Ab = np.zeros_like(Xb)
        ne = len(self.conn)
        Sb = np.zeros((ne, 4, 3, 3)); evb = np.zeros((ne, 4)); pb = np.zeros(ne); Tb = np.zeros((ne, 4))
        thb = np.zeros(3)
        tb = 0.0; dtb_carry = 0.0; tsb = ltsb = Cb = Lb = 0.0
        for k in range(nsteps - 1, -1, -1):
            (x, v, a, du, S, evp, p, Temp), sc, a_new = hist[k]
            dt = sc['dt']
            dtb = tb + dtb_carry  # t_{k+1} = t_k + dt ; dt carried from next step if unchanged
            xb, vb, ab, fb, dtb_n = self.nodal_backward(v, a, a_new, dt, Xb, Ub, Vb, Ab)
            dtb += dtb_n
            fbar = fb[self.conn]
            xb2, dub, Sb, evb, pb, thb_e, Tb = self.elem_backward(x, du, S, evp, p, th, fbar, Sb, evb, pb, Temp, Tb)
            thb += thb_e
            Xb = xb + xb2; Ub = dub; Vb = vb; Ab = ab
            # time-step computation of this increment
            dtb_carry = 0.0
            if k == 0:
                pass
            elif sc['rec']:
                dtb_tot = dtb + Cb
                critb = dtb_tot * self.safety * self.om / 2
                e = sc['amin']
                g = self.courant_grad(x, e)
                np.add.at(Xb, self.conn[e], critb * g)
                tb_add = tsb
                tsb, Cb, ltsb, Lb = ltsb, Lb, 0.0, 0.0
                tb += tb_add
            elif sc['ts'] != sc['lts']:
                t, ts, lts, C, L = sc['t'], sc['ts'], sc['lts'], sc['C'], sc['L']
                dx = (t - lts) / (ts - lts)
                Cb += dtb * dx; Lb += dtb * (1 - dx)
                dxb = dtb * (C - L)
                tb += dxb / (ts - lts)
                tsb += -dxb * (t - lts) / (ts - lts) ** 2
                ltsb += dxb * (t - ts) / (ts - lts) ** 2
            else:
                dtb_carry = dtb
        return xN, thb
