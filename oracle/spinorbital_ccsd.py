"""Independent spin-orbital CCSD (oracle cross-check; test infrastructure only).

Textbook Stanton-Gauss form (J. F. Stanton, J. Gauss, J. D. Watts,
R. J. Bartlett, J. Chem. Phys. 94, 4334 (1991)), valid for a non-canonical
Fock operator (its diagonal is moved into the denominators).  It shares no
code or intermediates with ``oracle.rhfccsd`` and is used on tiny closed-shell
systems to check the spin-adapted restatement of rhfccsd.cxx:194-327: for a
closed shell both must converge to the same energy and the same amplitudes
(t_ai = t[i alpha, a alpha]; T2[a,b,i,j] = t[i alpha, j beta, a alpha, b beta]).
The reference's own spin-orbital task (src/cc/ccsd.cxx:92-193) solves the same
equations through its spin-orbital tensor layer.
"""

import numpy as np


def spin_orbital_system(f, eri, o):
    """Spin-orbital Fock F[P,Q] and antisymmetrised <PQ||RS> from spatial f
    (n x n) and Mulliken eri[p,q,r,s].  Spin orbitals: occupied (i,s) ->
    2i+s, virtual (a,s) -> 2o+2a+s."""
    n = f.shape[0]
    v = n - o
    sp = np.concatenate([np.repeat(np.arange(o), 2), o + np.repeat(np.arange(v), 2)])
    sg = np.tile([0, 1], n)
    same = (sg[:, None] == sg[None, :]).astype(float)
    F = f[np.ix_(sp, sp)] * same
    # <PQ|RS> = (pr|qs) d(sP,sR) d(sQ,sS)
    g = eri[np.ix_(sp, sp, sp, sp)].transpose(0, 2, 1, 3)
    g = g * same[:, None, :, None] * same[None, :, None, :]
    return F, g - g.transpose(0, 1, 3, 2)


class SpinOrbitalCCSD:
    def __init__(self, F, W, nocc):
        self.no = nocc
        o, v = slice(0, nocc), slice(nocc, F.shape[0])
        self.f_oo, self.f_vv, self.f_ov = F[o, o], F[v, v], F[o, v]
        self.oooo, self.ooov = W[o, o, o, o], W[o, o, o, v]
        self.oovv, self.ovvo = W[o, o, v, v], W[o, v, v, o]
        self.ovov, self.vvvv = W[o, v, o, v], W[v, v, v, v]
        self.vvvo, self.ovoo = W[v, v, v, o], W[o, v, o, o]
        self.oovo, self.ovvv = W[o, o, v, o], W[o, v, v, v]
        self.vovv = W[v, o, v, v]
        self.vvoo = W[v, v, o, o]
        fo, fv = np.diag(self.f_oo), np.diag(self.f_vv)
        self.Dia = fo[:, None] - fv[None, :]
        self.Dijab = fo[:, None, None, None] + fo[None, :, None, None] - fv[None, None, :, None] - fv[None, None, None, :]
        self.t1 = self.f_ov / self.Dia
        self.t2 = self.oovv / self.Dijab

    def energy(self, t1=None, t2=None):
        t1 = self.t1 if t1 is None else t1
        t2 = self.t2 if t2 is None else t2
        e = np.einsum("ia,ia->", self.f_ov, t1)
        e += 0.25 * np.einsum("ijab,ijab->", self.oovv, t2)
        e += 0.5 * np.einsum("ijab,ia,jb->", self.oovv, t1, t1)
        return float(e)

    def update(self):
        t1, t2 = self.t1, self.t2
        es = lambda s, *a: np.einsum(s, *a, optimize=True)
        nv = self.f_vv.shape[0]
        no = self.no
        tau_t = t2 + 0.5 * (es("ia,jb->ijab", t1, t1) - es("ja,ib->ijab", t1, t1))
        tau = t2 + es("ia,jb->ijab", t1, t1) - es("ja,ib->ijab", t1, t1)
        fvv = self.f_vv - np.diag(np.diag(self.f_vv))
        foo = self.f_oo - np.diag(np.diag(self.f_oo))

        Fae = fvv - 0.5 * es("me,ma->ae", self.f_ov, t1) + es("mf,mafe->ae", t1, self.ovvv) \
            - 0.5 * es("mnaf,mnef->ae", tau_t, self.oovv)
        Fmi = foo + 0.5 * es("ie,me->mi", t1, self.f_ov) + es("ne,mnie->mi", t1, self.ooov) \
            + 0.5 * es("inef,mnef->mi", tau_t, self.oovv)
        Fme = self.f_ov + es("nf,mnef->me", t1, self.oovv)

        Wmnij = self.oooo + es("je,mnie->mnij", t1, self.ooov) - es("ie,mnje->mnij", t1, self.ooov) \
            + 0.25 * es("ijef,mnef->mnij", tau, self.oovv)
        Wabef = self.vvvv - es("mb,amef->abef", t1, self.vovv) + es("ma,bmef->abef", t1, self.vovv) \
            + 0.25 * es("mnab,mnef->abef", tau, self.oovv)
        Wmbej = self.ovvo + es("jf,mbef->mbej", t1, self.ovvv) - es("nb,mnej->mbej", t1, self.oovo) \
            - es("jnfb,mnef->mbej", 0.5 * t2 + es("jf,nb->jnfb", t1, t1), self.oovv)

        r1 = self.f_ov.copy()
        r1 += es("ie,ae->ia", t1, Fae)
        r1 -= es("ma,mi->ia", t1, Fmi)
        r1 += es("imae,me->ia", t2, Fme)
        r1 -= es("nf,naif->ia", t1, self.ovov)
        r1 -= 0.5 * es("imef,maef->ia", t2, self.ovvv)
        r1 -= 0.5 * es("mnae,nmei->ia", t2, self.oovo)

        r2 = self.oovv.copy()
        X = Fae - 0.5 * es("mb,me->be", t1, Fme)       # X[b,e]
        P = es("ijae,be->ijab", t2, X)
        r2 += P - P.transpose(0, 1, 3, 2)
        Y = Fmi + 0.5 * es("je,me->mj", t1, Fme)       # Y[m,j]
        P = es("imab,mj->ijab", t2, Y)
        r2 -= P - P.transpose(1, 0, 2, 3)
        r2 += 0.5 * es("mnab,mnij->ijab", tau, Wmnij)
        r2 += 0.5 * es("ijef,abef->ijab", tau, Wabef)
        P = es("imae,mbej->ijab", t2, Wmbej) - es("ie,ma,mbej->ijab", t1, t1, self.ovvo)
        r2 += P - P.transpose(1, 0, 2, 3) - P.transpose(0, 1, 3, 2) + P.transpose(1, 0, 3, 2)
        P = es("ie,abej->ijab", t1, self.vvvo)
        r2 += P - P.transpose(1, 0, 2, 3)
        P = es("ma,mbij->ijab", t1, self.ovoo)
        r2 -= P - P.transpose(0, 1, 3, 2)
        del nv, no
        return r1 / self.Dia, r2 / self.Dijab

    def solve(self, tol=1e-12, maxiter=200, ndiis=8):
        hist_t, hist_e = [], []
        e_old = self.energy()
        for it in range(1, maxiter + 1):
            t1n, t2n = self.update()
            err = np.concatenate([(t1n - self.t1).ravel(), (t2n - self.t2).ravel()])
            vec = np.concatenate([t1n.ravel(), t2n.ravel()])
            hist_t.append(vec)
            hist_e.append(err)
            if len(hist_t) > ndiis:
                hist_t.pop(0)
                hist_e.pop(0)
            if len(hist_t) > 2:
                k = len(hist_t)
                Bm = -np.ones((k + 1, k + 1))
                Bm[k, k] = 0.0
                for a in range(k):
                    for b in range(k):
                        Bm[a, b] = hist_e[a] @ hist_e[b]
                rhs = np.zeros(k + 1)
                rhs[k] = -1.0
                c = np.linalg.solve(Bm, rhs)[:k]
                vec = sum(ci * ti for ci, ti in zip(c, hist_t))
            n1 = self.t1.size
            self.t1 = vec[:n1].reshape(self.t1.shape)
            self.t2 = vec[n1:].reshape(self.t2.shape)
            e = self.energy()
            if np.max(np.abs(err)) < tol and abs(e - e_old) < tol:
                return e, it
            e_old = e
        raise RuntimeError("spin-orbital CCSD did not converge")
