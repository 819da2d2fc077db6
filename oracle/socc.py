"""Spin-orbital CCSD on a UHF reference, restated on numpy (oracle; test
infrastructure only).

Follows the reference's spin-orbital task statement by statement:
  * CCSD<U>::run / iterate      src/cc/ccsd.cxx:30-193
  * task spec / defaults        src/cc/ccsd.cxx (convergence 1e-9, 50
                                iterations, DIIS order 5, start 1)
  * the spin-orbital integrals  src/operator/aomoints.cxx (antisymmetrised
                                <PQ||RS> over the UHF spin orbitals)
Spin orbitals are dense here: occupied alpha, occupied beta, virtual alpha,
virtual beta.  The reference's tensors are antisymmetric; a product written
into an antisymmetric output is antisymmetrised over every index pair whose
two indices come from different operands (the implicit P(pq) of its
spin-orbital contractions), which is what `mult` does below.

Pinned by the reference's CH2 triplet known answers (test/input.aq:67-70:
UHF SCF, MP2, CCD, CCSD to 1e-9) and, on closed shells, by equality with the
spin-adapted restatement (oracle.rhfccsd) and the independent Stanton-Gauss
code (oracle.spinorbital_ccsd).
"""

import numpy as np

from .rhfccsd import DIIS

# antisymmetric index pairs of each tensor kind (by position)
AS_PAIRS = {
    "abij": [(0, 1), (2, 3)], "ijab": [(0, 1), (2, 3)], "abcd": [(0, 1), (2, 3)], "ijkl": [(0, 1), (2, 3)],
    "amef": [(2, 3)], "abej": [(0, 1)], "mnej": [(0, 1)], "amij": [(2, 3)],
}


def spin_orbitals(Ca, Cb, Fa, Fb, eri_ao, na, nb):
    """Dense spin-orbital Fock f[P,Q] and antisymmetrised g[P,Q,R,S] = <PQ||RS>,
    with the orbital order occ-alpha, occ-beta, vir-alpha, vir-beta."""
    n = Ca.shape[1]
    cols = [Ca[:, :na], Cb[:, :nb], Ca[:, na:], Cb[:, nb:]]
    spin = np.concatenate([np.zeros(na), np.ones(nb), np.zeros(n - na), np.ones(n - nb)]).astype(int)
    C = np.concatenate(cols, axis=1)
    Fs = [Fa, Fb]
    f = np.zeros((C.shape[1],) * 2)
    for s in (0, 1):
        idx = np.where(spin == s)[0]
        f[np.ix_(idx, idx)] = C[:, idx].T @ Fs[s] @ C[:, idx]
    mo = np.einsum("mnls,mp->pnls", eri_ao, C, optimize=True)
    mo = np.einsum("pnls,nq->pqls", mo, C, optimize=True)
    mo = np.einsum("pqls,lr->pqrs", mo, C, optimize=True)
    mo = np.einsum("pqrs,st->pqrt", mo, C, optimize=True)
    same = (spin[:, None] == spin[None, :]).astype(float)
    mo *= same[:, :, None, None] * same[None, None, :, :]
    dirac = mo.transpose(0, 2, 1, 3)                      # <PQ|RS> = (PR|QS)
    return f, dirac - dirac.transpose(0, 1, 3, 2), spin


def mult(alpha, A, ia, B, ib, beta, C, ic, pairs):
    """C[ic] = beta C + alpha * AS(A[ia] B[ib]) with the implicit antisymmetriser
    over the output pairs whose indices come from different operands."""
    X = alpha * np.einsum(f"{ia},{ib}->{ic}", A, B, optimize=True)
    for p, q in pairs:
        a, b = ic[p], ic[q]
        if (a in ia) != (b in ia) or (a in ib) != (b in ib):
            perm = list(range(len(ic)))
            perm[p], perm[q] = q, p
            X = X - X.transpose(perm)
    C *= beta
    C += X


class SpinOrbitalCCSD:
    """ccsd.cxx iterate on dense spin-orbital tensors."""

    def __init__(self, f, g, nocc, config=None, freeze_t1=False):
        from .rhfccsd import default_config
        self.config = default_config() if config is None else config
        o = nocc
        O, V = slice(0, o), slice(o, f.shape[0])
        self.no, self.nv = o, f.shape[0] - o
        self.fAI, self.fME, self.fAE, self.fMI = f[V, O], f[O, V], f[V, V], f[O, O]
        self.VABIJ, self.VMNEF = g[V, V, O, O], g[O, O, V, V]
        self.VAMEF, self.VABEJ = g[V, O, V, V], g[V, V, V, O]
        self.VABEF, self.VMNIJ = g[V, V, V, V], g[O, O, O, O]
        self.VMNEJ, self.VAMIJ, self.VAMEI = g[O, O, V, O], g[V, O, O, O], g[V, O, V, O]
        self.freeze_t1 = freeze_t1
        dA, dI = -np.diag(self.fAE), np.diag(self.fMI)
        self.D1 = dA[:, None] + dI[None, :]
        self.D2 = dA[:, None, None, None] + dA[None, :, None, None] + dI[None, None, :, None] + dI[None, None, None, :]
        # T(1) = fAI, T(2) = VABIJ, T.weight(D)   (ccsd.cxx:55-60)
        self.T1 = self._weight(self.fAI.copy(), self.D1)
        if freeze_t1:
            self.T1[...] = 0.0
        self.T2 = self._weight(self.VABIJ.copy(), self.D2)
        self.Tau = np.zeros_like(self.T2)
        self._tau()
        self.mp2 = float(np.sum(self.fAI * self.T1) + 0.25 * np.sum(self.VABIJ * self.Tau))   # :65
        self.diis = DIIS(self.config["diis"], 2, 2)
        self.energy = 0.0
        self.conv = np.finfo(np.float64).max
        self.iter = 0

    @staticmethod
    def _weight(X, D):
        den = D.copy()
        small = np.abs(den) < 1e-4
        den[small] = 1.0
        X = X / den
        X[small] = 0.0
        return X

    def _tau(self):
        self.Tau[...] = self.T2
        mult(0.5, self.T1, "ai", self.T1, "bj", 1.0, self.Tau, "abij", AS_PAIRS["abij"])

    def iterate(self):
        T1, T2, Tau = self.T1, self.T2, self.Tau
        no, nv = self.no, self.nv
        self._tau()                                                                   # :114-115
        FME = self.fME.copy()                                                         # :121
        mult(1.0, self.VMNEF, "mnef", T1, "fn", 1.0, FME, "me", [])                   # :122
        FMI = self.fMI.copy()                                                         # :124
        mult(0.5, self.VMNEF, "mnef", T2, "efin", 1.0, FMI, "mi", [])                 # :125
        mult(1.0, FME, "me", T1, "ei", 1.0, FMI, "mi", [])                            # :126
        mult(1.0, self.VMNEJ, "nmfi", T1, "fn", 1.0, FMI, "mi", [])                   # :127
        FAE = self.fAE.copy()                                                         # :129
        mult(-0.5, self.VMNEF, "mnef", T2, "afmn", 1.0, FAE, "ae", [])                # :130
        mult(-1.0, FME, "me", T1, "am", 1.0, FAE, "ae", [])                           # :131
        mult(1.0, self.VAMEF, "amef", T1, "fm", 1.0, FAE, "ae", [])                   # :132
        WMNIJ = self.VMNIJ.copy()                                                     # :134
        mult(0.5, self.VMNEF, "mnef", Tau, "efij", 1.0, WMNIJ, "mnij", AS_PAIRS["ijkl"])   # :135
        mult(1.0, self.VMNEJ, "mnej", T1, "ei", 1.0, WMNIJ, "mnij", AS_PAIRS["ijkl"])     # :136
        WMNEJ = self.VMNEJ.copy()                                                     # :138
        mult(1.0, self.VMNEF, "mnef", T1, "fj", 1.0, WMNEJ, "mnej", AS_PAIRS["mnej"])     # :139
        WAMIJ = self.VAMIJ.copy()                                                     # :141
        mult(0.5, self.VAMEF, "amef", Tau, "efij", 1.0, WAMIJ, "amij", AS_PAIRS["amij"])  # :142
        mult(1.0, self.VAMEI, "amej", T1, "ei", 1.0, WAMIJ, "amij", AS_PAIRS["amij"])     # :143
        WAMEI = self.VAMEI.copy()                                                     # :145
        mult(0.5, self.VMNEF, "mnef", T2, "afni", 1.0, WAMEI, "amei", [])             # :146
        mult(1.0, self.VAMEF, "amef", T1, "fi", 1.0, WAMEI, "amei", [])               # :147
        mult(-1.0, WMNEJ, "nmei", T1, "an", 1.0, WAMEI, "amei", [])                   # :148

        Z1 = self.fAI.copy()                                                          # :157
        mult(1.0, self.fAE, "ae", T1, "ei", 1.0, Z1, "ai", [])                        # :158
        mult(-1.0, FMI, "mi", T1, "am", 1.0, Z1, "ai", [])                            # :159
        mult(-1.0, self.VAMEI, "amei", T1, "em", 1.0, Z1, "ai", [])                   # :160
        mult(1.0, FME, "me", T2, "aeim", 1.0, Z1, "ai", [])                           # :161
        mult(0.5, self.VAMEF, "amef", Tau, "efim", 1.0, Z1, "ai", [])                 # :162
        mult(-0.5, WMNEJ, "mnei", T2, "eamn", 1.0, Z1, "ai", [])                      # :163
        P = AS_PAIRS["abij"]
        Z2 = self.VABIJ.copy()                                                        # :165
        mult(1.0, self.VABEJ, "abej", T1, "ei", 1.0, Z2, "abij", P)                   # :166
        mult(-1.0, WAMIJ, "amij", T1, "bm", 1.0, Z2, "abij", P)                       # :167
        mult(1.0, FAE, "ae", T2, "ebij", 1.0, Z2, "abij", P)                          # :168
        mult(-1.0, FMI, "mi", T2, "abmj", 1.0, Z2, "abij", P)                         # :169
        mult(0.5, self.VABEF, "abef", Tau, "efij", 1.0, Z2, "abij", P)                # :170
        mult(0.5, WMNIJ, "mnij", Tau, "abmn", 1.0, Z2, "abij", P)                     # :171
        mult(1.0, WAMEI, "amei", T2, "ebjm", 1.0, Z2, "abij", P)                      # :172

        Z1 = self._weight(Z1, self.D1)                                                # :178
        Z2 = self._weight(Z2, self.D2)
        if self.freeze_t1:
            Z1[...] = 0.0
        T1 += Z1                                                                      # :179
        T2 += Z2
        self._tau()                                                                   # :181-182
        self.energy = float(np.sum(self.fAI * T1) + 0.25 * np.sum(self.VABIJ * self.Tau))   # :183
        self.conv = max(np.abs(Z1).max(), np.abs(Z2).max())                           # :184
        self.diis.extrapolate([T1, T2], [Z1, Z2])                                     # :186
        self.iter += 1

    def run(self, max_iterations=None):
        maxiter = self.config["max_iterations"] if max_iterations is None else max_iterations
        self.conv = np.finfo(np.float64).max
        it = 1
        while it <= maxiter and not self.conv < self.config["convergence"]:
            self.iterate()
            it += 1
        return self.conv < self.config["convergence"]
