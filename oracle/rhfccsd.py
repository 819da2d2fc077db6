"""Closed-shell (RHF) CCSD restated on numpy (oracle; test infrastructure only).

Follows, statement by statement, the reference's
  * RHFCCSD setup            src/cc/rhfccsd.cxx:45-104
  * RHFCCSD::iterate         src/cc/rhfccsd.cxx:194-327
  * task spec / defaults     src/cc/rhfccsd.cxx:332-353
  * DIIS::extrapolate        src/convergence/diis.hpp:16-227
  * Iterative::run           src/util/iterative.hpp:126-188
  * Hbar products            src/cc/rhfccsd.cxx:108-188 (RHFCCSD.hbar; pinned by
                             the T = 0 limit, FME = dE/dT1 / 2, and through their
                             consumer: the EOM-EE-CCSD singlet sigma vector built
                             from them, src/cc/rhfeomeeccsd.cxx:379-424
                             (eom_singlet_sigma), equals the exact derivative of
                             the CC residual along any R at converged T)
  * Denominator              src/operator/denominator.hpp:21-89
with the operator semantics of oracle.ctf_ops (CTFTensor, ctf_tensor.cxx:403-514).

Inputs use the Dirac block layout the reference's AO->MO step produces
(rhfaomoints.cxx:58-65,168; aomoints.cxx:806):
  VABCD[a,b,c,d] = <ab|cd>, VABCI[a,b,c,i], VABIJ[a,b,i,j], VAIBJ[a,i,b,j],
  VAIJB[a,i,j,b], VAIJK[a,i,j,k], VIJKL[i,j,k,l];
Fock blocks fAI[a,i], fAB[a,b], fIJ[i,j].
"""

import numpy as np
import scipy.linalg.lapack as lapack

from . import ctf_ops as ops

mult = ops.contract
tsum = ops.tsum

BLOCKS = ("VABCD", "VABCI", "VABIJ", "VAIBJ", "VAIJB", "VAIJK", "VIJKL")


def default_config():
    """Task spec defaults (rhfccsd.cxx:332-353)."""
    return {"convergence": 1e-9, "max_iterations": 50, "conv_type": "MAXE",
            "diis": {"damping": 0.0, "start": 1, "order": 5, "jacobi": False},
            # not a reference option: True holds T1 at zero, which turns the
            # equations into CCD (the check against test/input.aq:22)
            "freeze_t1": False}


class Denominator:
    """dA = -f_AA, dI = +f_II (denominator.hpp:48-67); C1 means one irrep."""

    def __init__(self, fAB, fIJ):
        self.dA = -np.diag(fAB).copy()
        self.dI = np.diag(fIJ).copy()


def inner(a_list, b_list):
    """DefaultInnerProd: sum over components of scalar(a*b) (diis.hpp:17-32)."""
    p = 0.0
    for a, b in zip(a_list, b_list):
        p += float(np.vdot(a, b))
    return p


class DIIS:
    """Pulay extrapolation restating diis.hpp:36-227 (nx = ndx = 2 for T1,T2)."""

    def __init__(self, config, nx=1, ndx=1):
        self.nextrap = int(config["order"])
        self.start = int(config["start"])
        self.damping = float(config["damping"])
        self.nx, self.ndx = nx, ndx
        n = self.nextrap
        self.e = np.zeros((n + 1, n + 1))
        self.c = np.zeros(n + 1)
        self.old_x = [None] * n
        self.old_dx = [None] * n
        self.last_info = 0
        self.used_eps_retry = False

    def extrapolate(self, x, dx):
        """x, dx: lists of arrays, updated in place (diis.hpp:70-227)."""
        assert len(x) == self.nx and len(dx) == self.ndx
        n = self.nextrap
        if n <= 1:
            return
        # slot k holds iteration n-k (diis.hpp:82-85; marray rotate(-1) is a
        # right rotation by one, as std::rotate(begin, end-1, end))
        self.old_x = [self.old_x[-1]] + self.old_x[:-1]
        self.old_dx = [self.old_dx[-1]] + self.old_dx[:-1]
        self.e = np.roll(self.e, (1, 1), axis=(0, 1))
        self.c = np.roll(self.c, 1)

        self.old_x[0] = [xi.copy() for xi in x]         # :92-105
        self.old_dx[0] = [di.copy() for di in dx]       # :107-120

        e, c = self.e, self.c
        e[0, 0] = inner(dx, dx)                         # :122
        nreal = 1
        i = 1
        while i < n and self.old_dx[i] is not None:     # :129-135
            e[i, 0] = inner(dx, self.old_dx[i])
            e[0, i] = e[i, 0]
            nreal += 1
            i += 1

        for i in range(nreal):                          # :141-149
            e[i, nreal] = -1.0
            e[nreal, i] = -1.0
            c[i] = 0.0
        e[nreal, nreal] = 0.0
        c[nreal] = -1.0

        if nreal == 1:                                  # :151
            return

        self.start -= 1
        if self.start > 1:                              # :153-164
            if self.damping > 0.0:
                for i in range(self.nx):
                    self.old_x[0][i] = (self.damping - 1) * self.old_x[0][i] + self.damping * self.old_x[1][i]
            return

        m = nreal + 1                                   # :166-188
        A = e[:m, :m].copy()
        _, _, sol, info = lapack.dsysv(A, c[:m].reshape(m, 1).copy(), lower=0)
        self.used_eps_retry = False
        if info > 0:
            eps = e[nreal - 1, nreal - 1] * np.finfo(np.float64).eps
            A = e[:m, :m].copy()
            A[np.arange(nreal), np.arange(nreal)] += eps
            _, _, sol, info = lapack.dsysv(A, c[:m].reshape(m, 1).copy(), lower=0)
            self.used_eps_retry = True
        self.last_info = info
        if info != 0:
            raise RuntimeError(f"DIIS: Info in hesv: {info}")
        c[:m] = sol[:, 0]

        for j in range(self.ndx):                       # :205-226
            dx[j][...] = self.old_dx[0][j] * c[0]
        for j in range(self.nx):
            x[j][...] = self.old_x[0][j] * c[0]
        for i in range(1, nreal):
            for j in range(self.ndx):
                dx[j] += self.old_dx[i][j] * c[i]
            for j in range(self.nx):
                x[j] += self.old_x[i][j] * c[i]


class RHFCCSD:
    """Spin-adapted closed-shell CCSD (rhfccsd.cxx).  ``f`` holds fAI/fAB/fIJ,
    ``V`` the seven Dirac blocks."""

    def __init__(self, f, V, config=None):
        self.config = default_config() if config is None else config
        self.fAI = np.asarray(f["fAI"], dtype=np.float64)
        self.fAB = np.asarray(f["fAB"], dtype=np.float64)
        self.fIJ = np.asarray(f["fIJ"], dtype=np.float64)
        self.V = {k: np.asarray(V[k], dtype=np.float64) for k in BLOCKS}
        nv, no = self.fAI.shape
        self.no, self.nv = no, nv
        self.energy = 0.0
        self.conv = np.finfo(np.float64).max
        self.iter = 0
        self.history = []          # (energy, conv) per iteration
        self._setup()

    # rhfccsd.cxx:45-104
    def _setup(self):
        no, nv = self.no, self.nv
        V = self.V
        self.D = Denominator(self.fAB, self.fIJ)
        dA, dI = self.D.dA, self.D.dI
        self.T1 = np.zeros((nv, no))
        self.T2 = np.zeros((nv, nv, no, no))
        self.Z1 = np.zeros((nv, no))
        self.Z2 = np.zeros((nv, nv, no, no))
        self.Tau = np.zeros((nv, nv, no, no))
        self.T2SA = np.zeros((nv, nv, no, no))
        self.TauSA = np.zeros((nv, nv, no, no))
        self.VABIJSA = np.zeros((nv, nv, no, no))
        self.VAIJBSA = np.zeros((nv, no, no, nv))
        self.FAE = np.zeros((nv, nv))
        self.FMI = np.zeros((no, no))
        self.FME = np.zeros((no, nv))
        self.WMNIJ = np.zeros((no, no, no, no))
        self.WMNEJ = np.zeros((no, no, nv, no))
        self.WAMIJ = np.zeros((nv, no, no, no))
        self.WAMEI = np.zeros((nv, no, nv, no))
        self.WAMIE = np.zeros((nv, no, no, nv))

        self.T1[...] = self.fAI                                   # :79
        ops.weight(self.T1, [dA, dI])                             # :80
        if self.config.get("freeze_t1"):
            self.T1[...] = 0.0      # CCD: the CCSD equations restricted to T1 = 0
        self.T2[...] = V["VABIJ"]                                 # :82
        ops.weight(self.T2, [dA, dA, dI, dI])                     # :83

        tsum(1.0, self.T2, "abij", 0.0, self.Tau, "abij")          # :85
        mult(1.0, self.T1, "ai", self.T1, "bj", 1.0, self.Tau, "abij")  # :86

        tsum(2.0, V["VABIJ"], "abij", 0.0, self.VABIJSA, "abij")  # :88
        tsum(-1.0, V["VABIJ"], "abji", 1.0, self.VABIJSA, "abij")  # :89

        tsum(2.0, V["VAIJB"], "aijb", 0.0, self.VAIJBSA, "aijb")  # :91
        tsum(-1.0, V["VAIBJ"], "aibj", 1.0, self.VAIJBSA, "aijb")  # :92

        self.mp2 = 2 * ops.dot(self.fAI, "ai", self.T1, "ai") + \
            ops.dot(self.VABIJSA, "abij", self.Tau, "abij")        # :94
        self.diis = DIIS(self.config["diis"], 2, 2)                # :98

    # rhfccsd.cxx:194-327
    def iterate(self):
        self._residual()
        T1, T2, Z1, Z2, Tau = self.T1, self.T2, self.Z1, self.Z2, self.Tau
        fAI, VABIJSA = self.fAI, self.VABIJSA
        dA, dI = self.D.dA, self.D.dI
        ops.weight(Z1, [dA, dI])                                          # :313
        ops.weight(Z2, [dA, dA, dI, dI])                                  # :314
        if self.config.get("freeze_t1"):
            Z1[...] = 0.0
        T1 += Z1                                                          # :316
        T2 += Z2                                                          # :317

        tsum(1.0, T2, "abij", 0.0, Tau, "abij")                           # :319
        mult(1.0, T1, "ai", T1, "bj", 1.0, Tau, "abij")                    # :320

        self.energy = 2 * ops.dot(fAI, "ai", T1, "ai") + ops.dot(VABIJSA, "abij", Tau, "abij")  # :322
        self.conv = max(ops.norm(Z1, 0), ops.norm(Z2, 0))                 # :323
        self.diis.extrapolate([T1, T2], [Z1, Z2])                         # :325-326

    def residual_at(self, T1, T2):
        """The unweighted residual (Z1, Z2) of rhfccsd.cxx:231-309 at the given
        amplitudes; the solver's own state is left as it was."""
        names = ("T1", "T2", "Z1", "Z2", "Tau", "T2SA", "TauSA", "FAE", "FMI", "FME", "WMNIJ", "WMNEJ",
                 "WAMIJ", "WAMEI", "WAMIE")
        saved = {k: getattr(self, k).copy() for k in names}
        try:
            self.T1[...] = T1
            self.T2[...] = T2
            self._residual()
            return self.Z1.copy(), self.Z2.copy()
        finally:
            for k, x in saved.items():
                getattr(self, k)[...] = x

    # rhfccsd.cxx:231-309: intermediates and the residual Z1, Z2 of the current T
    def _residual(self):
        V = self.V
        fAI, fAB, fIJ = self.fAI, self.fAB, self.fIJ
        VABCD, VABCI, VABIJ = V["VABCD"], V["VABCI"], V["VABIJ"]
        VAIBJ, VAIJB, VAIJK, VIJKL = V["VAIBJ"], V["VAIJB"], V["VAIJK"], V["VIJKL"]
        T1, T2, Z1, Z2 = self.T1, self.T2, self.Z1, self.Z2
        Tau, T2SA, TauSA = self.Tau, self.T2SA, self.TauSA
        VABIJSA, VAIJBSA = self.VABIJSA, self.VAIJBSA
        FAE, FMI, FME = self.FAE, self.FMI, self.FME
        WMNIJ, WMNEJ, WAMIJ = self.WMNIJ, self.WMNEJ, self.WAMIJ
        WAMEI, WAMIE = self.WAMEI, self.WAMIE
        dA, dI = self.D.dA, self.D.dI

        tsum(1.0, T2, "abij", 0.0, Tau, "abij")                           # :231
        mult(1.0, T1, "ai", T1, "bj", 1.0, Tau, "abij")                    # :232
        tsum(2.0, Tau, "abij", 0.0, TauSA, "abij")                        # :234
        tsum(-1.0, Tau, "abji", 1.0, TauSA, "abij")                        # :235
        tsum(2.0, T2, "abij", 0.0, T2SA, "abij")                          # :237
        tsum(-1.0, T2, "abji", 1.0, T2SA, "abij")                          # :238

        tsum(1.0, fAI, "em", 0.0, FME, "me")                              # :244
        mult(1.0, VABIJSA, "efmn", T1, "fn", 1.0, FME, "me")              # :245

        tsum(1.0, fIJ, "mi", 0.0, FMI, "mi")                              # :247
        mult(1.0, VABIJSA, "efmn", T2, "efin", 1.0, FMI, "mi")            # :248
        mult(1.0, FME, "me", T1, "ei", 1.0, FMI, "mi")                    # :249
        mult(2.0, VAIJK, "finm", T1, "fn", 1.0, FMI, "mi")                # :250
        mult(-1.0, VAIJK, "fimn", T1, "fn", 1.0, FMI, "mi")               # :251

        tsum(1.0, fAB, "ae", 0.0, FAE, "ae")                              # :253
        mult(-1.0, VABIJSA, "efmn", T2, "afmn", 1.0, FAE, "ae")           # :254
        mult(-1.0, FME, "me", T1, "am", 1.0, FAE, "ae")                   # :255
        mult(2.0, VABCI, "efam", T1, "fm", 1.0, FAE, "ae")                # :256
        mult(-1.0, VABCI, "feam", T1, "fm", 1.0, FAE, "ae")               # :257

        tsum(0.5, VIJKL, "mnij", 0.0, WMNIJ, "mnij")                      # :259
        mult(0.5, VABIJ, "efmn", Tau, "efij", 1.0, WMNIJ, "mnij")         # :260
        mult(1.0, VAIJK, "ejmn", T1, "ei", 1.0, WMNIJ, "mnij")            # :261
        tsum(1.0, WMNIJ, "nmji", 1.0, WMNIJ, "mnij")                      # :262 (aliased)

        tsum(1.0, VAIJK, "ejmn", 0.0, WMNEJ, "mnej")                      # :264
        mult(1.0, VABIJ, "efmn", T1, "fj", 1.0, WMNEJ, "mnej")            # :265

        tsum(1.0, VAIJK, "amij", 0.0, WAMIJ, "amij")                      # :267
        mult(1.0, VABCI, "efam", Tau, "efij", 1.0, WAMIJ, "amij")         # :268
        mult(1.0, VAIBJ, "amej", T1, "ei", 1.0, WAMIJ, "amij")            # :269
        mult(1.0, VAIJB, "amie", T1, "ej", 1.0, WAMIJ, "amij")            # :270

        tsum(1.0, VAIJBSA, "amie", 0.0, WAMIE, "amie")                    # :272
        mult(0.5, VABIJSA, "efmn", T2SA, "afin", 1.0, WAMIE, "amie")      # :273
        mult(2.0, VABCI, "feam", T1, "fi", 1.0, WAMIE, "amie")            # :274
        mult(-1.0, VABCI, "efam", T1, "fi", 1.0, WAMIE, "amie")           # :275
        mult(-2.0, WMNEJ, "mnei", T1, "an", 1.0, WAMIE, "amie")           # :276
        mult(1.0, WMNEJ, "nmei", T1, "an", 1.0, WAMIE, "amie")            # :277

        tsum(1.0, VAIBJ, "amei", 0.0, WAMEI, "amei")                      # :279
        mult(-0.5, VABIJ, "efnm", T2, "afni", 1.0, WAMEI, "amei")         # :280
        mult(1.0, VABCI, "efam", T1, "fi", 1.0, WAMEI, "amei")            # :281
        mult(-1.0, WMNEJ, "nmei", T1, "an", 1.0, WAMEI, "amei")           # :282

        tsum(1.0, fAI, "ai", 0.0, Z1, "ai")                               # :290
        mult(1.0, fAB, "ae", T1, "ei", 1.0, Z1, "ai")                     # :291
        mult(-1.0, FMI, "mi", T1, "am", 1.0, Z1, "ai")                    # :292
        mult(1.0, VAIJBSA, "amie", T1, "em", 1.0, Z1, "ai")               # :293
        mult(1.0, FME, "me", T2SA, "aeim", 1.0, Z1, "ai")                 # :294
        mult(1.0, VABCI, "efam", TauSA, "efim", 1.0, Z1, "ai")            # :295
        mult(-1.0, WMNEJ, "mnei", T2SA, "eamn", 1.0, Z1, "ai")            # :296

        tsum(0.5, VABIJ, "abij", 0.0, Z2, "abij")                         # :298
        mult(1.0, VABCI, "abej", T1, "ei", 1.0, Z2, "abij")               # :299
        mult(-1.0, WAMIJ, "amij", T1, "bm", 1.0, Z2, "abij")              # :300
        mult(1.0, FAE, "ae", T2, "ebij", 1.0, Z2, "abij")                 # :301
        mult(-1.0, FMI, "mi", T2, "abmj", 1.0, Z2, "abij")                # :302
        mult(0.5, VABCD, "abef", Tau, "efij", 1.0, Z2, "abij")            # :303
        mult(0.5, WMNIJ, "mnij", Tau, "abmn", 1.0, Z2, "abij")            # :304
        mult(0.5, WAMIE, "amie", T2SA, "ebmj", 1.0, Z2, "abij")           # :305
        mult(-0.5, WAMEI, "amei", T2, "ebjm", 1.0, Z2, "abij")            # :306
        mult(-1.0, WAMEI, "amej", T2, "ebim", 1.0, Z2, "abij")            # :307
        tsum(1.0, Z2, "baji", 1.0, Z2, "abij")                            # :309 (aliased)

    # rhfccsd.cxx:108-188
    def hbar(self, with_wabej=False):
        """Hbar products of the current amplitudes, in place (FAE ... WAMIE;
        WABEJ [v,v,v,o] when asked)."""
        V = self.V
        fAI, fAB, fIJ = self.fAI, self.fAB, self.fIJ
        VABCD, VABCI, VABIJ = V["VABCD"], V["VABCI"], V["VABIJ"]
        VAIBJ, VAIJB, VAIJK, VIJKL = V["VAIBJ"], V["VAIJB"], V["VAIJK"], V["VIJKL"]
        T1, T2, Tau, T2SA = self.T1, self.T2, self.Tau, self.T2SA
        VABIJSA, VAIJBSA = self.VABIJSA, self.VAIJBSA
        FAE, FMI, FME = self.FAE, self.FMI, self.FME
        WMNIJ, WMNEJ, WAMIJ = self.WMNIJ, self.WMNEJ, self.WAMIJ
        WAMEI, WAMIE = self.WAMEI, self.WAMIE

        tsum(1.0, T2, "abij", 0.0, Tau, "abij")                           # :119
        mult(1.0, T1, "ai", T1, "bj", 1.0, Tau, "abij")                    # :120
        tsum(2.0, T2, "abij", 0.0, T2SA, "abij")                          # :122
        tsum(-1.0, T2, "abji", 1.0, T2SA, "abij")                         # :123

        tsum(1.0, fAI, "em", 0.0, FME, "me")                              # :125
        mult(1.0, VABIJSA, "efmn", T1, "fn", 1.0, FME, "me")              # :126

        tsum(1.0, fIJ, "mi", 0.0, FMI, "mi")                              # :128
        mult(1.0, VABIJSA, "efmn", T2, "efin", 1.0, FMI, "mi")            # :129
        mult(1.0, FME, "me", T1, "ei", 1.0, FMI, "mi")                    # :130
        mult(2.0, VAIJK, "finm", T1, "fn", 1.0, FMI, "mi")                # :131
        mult(-1.0, VAIJK, "fimn", T1, "fn", 1.0, FMI, "mi")               # :132

        tsum(1.0, fAB, "ae", 0.0, FAE, "ae")                              # :134
        mult(-1.0, VABIJSA, "efmn", T2, "afmn", 1.0, FAE, "ae")           # :135
        mult(-1.0, FME, "me", T1, "am", 1.0, FAE, "ae")                   # :136
        mult(2.0, VABCI, "efam", T1, "fm", 1.0, FAE, "ae")                # :137
        mult(-1.0, VABCI, "feam", T1, "fm", 1.0, FAE, "ae")               # :138

        tsum(0.5, VIJKL, "mnij", 0.0, WMNIJ, "mnij")                      # :140
        mult(0.5, VABIJ, "efmn", Tau, "efij", 1.0, WMNIJ, "mnij")         # :141
        mult(1.0, VAIJK, "ejmn", T1, "ei", 1.0, WMNIJ, "mnij")            # :142
        tsum(1.0, WMNIJ, "nmji", 1.0, WMNIJ, "mnij")                      # :143 (aliased)

        tsum(1.0, VAIJK, "ejmn", 0.0, WMNEJ, "mnej")                      # :145
        mult(1.0, VABIJ, "efmn", T1, "fj", 1.0, WMNEJ, "mnej")            # :146

        tsum(1.0, VAIJK, "amij", 0.0, WAMIJ, "amij")                      # :148
        mult(1.0, VABCI, "efam", Tau, "efij", 1.0, WAMIJ, "amij")         # :149
        mult(1.0, VAIBJ, "amej", T1, "ei", 1.0, WAMIJ, "amij")            # :150
        mult(1.0, VAIJB, "amie", T1, "ej", 1.0, WAMIJ, "amij")            # :151
        mult(-1.0, WMNIJ, "nmij", T1, "an", 1.0, WAMIJ, "amij")           # :152
        mult(1.0, WMNEJ, "nmej", T2SA, "aein", 1.0, WAMIJ, "amij")        # :153
        mult(-1.0, WMNEJ, "mnej", T2, "aein", 1.0, WAMIJ, "amij")         # :154
        mult(-1.0, WMNEJ, "mnei", T2, "aenj", 1.0, WAMIJ, "amij")         # :155
        mult(1.0, FME, "me", T2, "aeij", 1.0, WAMIJ, "amij")              # :156

        tsum(1.0, VAIJBSA, "amie", 0.0, WAMIE, "amie")                    # :158
        mult(1.0, VABIJSA, "efmn", T2SA, "afin", 1.0, WAMIE, "amie")      # :159
        mult(2.0, VABCI, "feam", T1, "fi", 1.0, WAMIE, "amie")            # :160
        mult(-1.0, VABCI, "efam", T1, "fi", 1.0, WAMIE, "amie")           # :161

        tsum(1.0, VAIBJ, "amei", 0.0, WAMEI, "amei")                      # :163
        mult(-1.0, VABIJ, "efnm", T2, "afni", 1.0, WAMEI, "amei")         # :164
        mult(1.0, VABCI, "efam", T1, "fi", 1.0, WAMEI, "amei")            # :165

        if with_wabej:                                                    # :167-181
            nv, no = self.nv, self.no
            W = self.WABEJ = np.zeros((nv, nv, nv, no))
            tsum(1.0, VABCI, "abej", 0.0, W, "abej")
            mult(1.0, WMNEJ, "mnej", Tau, "abmn", 1.0, W, "abej")
            mult(-0.5, WAMIE, "bmje", T1, "am", 1.0, W, "abej")
            mult(-0.5, WAMEI, "bmej", T1, "am", 1.0, W, "abej")
            mult(-1.0, WAMEI, "amej", T1, "bm", 1.0, W, "abej")
            mult(1.0, VABCD, "abef", T1, "fj", 1.0, W, "abej")
            mult(1.0, VABCI, "efam", T2SA, "fbmj", 1.0, W, "abej")
            mult(-1.0, VABCI, "feam", T2, "fbmj", 1.0, W, "abej")
            mult(-1.0, VABCI, "febm", T2, "afmj", 1.0, W, "abej")
            mult(-1.0, FME, "me", T2, "abmj", 1.0, W, "abej")

        mult(-2.0, WMNEJ, "mnei", T1, "an", 1.0, WAMIE, "amie")           # :183
        mult(1.0, WMNEJ, "nmei", T1, "an", 1.0, WAMIE, "amie")            # :184

        mult(-1.0, WMNEJ, "nmei", T1, "an", 1.0, WAMEI, "amei")           # :186

    def converged(self):
        """Iterative::isConverged: every conv strictly below tol (iterative.hpp:178-188)."""
        return self.conv < self.config["convergence"]

    def run(self, max_iterations=None, callback=None):
        """Iterative::run (iterative.hpp:126-166).  Non-convergence is only
        logged by the reference; here it is reported through the return value
        (True = converged) while the state is kept, as there."""
        maxiter = self.config["max_iterations"] if max_iterations is None else max_iterations
        self.conv = np.finfo(np.float64).max
        self.iter = 1
        while self.iter <= maxiter and not self.converged():
            self.iterate()
            self.history.append((self.energy, self.conv))
            if callback is not None:
                callback(self)
            self.iter += 1
        return self.converged()


def eom_singlet_sigma(cc, R1, R2):
    """Singlet sigma vector of the EOM-EE-CCSD consumer of the Hbar products,
    restated from src/cc/rhfeomeeccsd.cxx:381-424 (the `else` branch, the
    singlet of RHFEOMEECCSD<U>::iterate), with the products of cc.hbar(True):
    ccsd.F(AE/MI/ME) = FAE/FMI/FME, ccsd.W(Am,Ie) = WAMIE (WAMIESA there),
    ccsd.W(Am,Ei) = WAMEI, ccsd.W(Mn,Ej) = WMNEJ, ccsd.W(Ab,Ej) = WABEJ,
    ccsd.W(Am,Ij) = WAMIJ, ccsd.W(Mn,Ij) = WMNIJ.  Returns (Z1, Z2)."""
    V = cc.V
    VABCD, VABCI, VABIJ = V["VABCD"], V["VABCI"], V["VABIJ"]
    T1, T2 = cc.T1, cc.T2
    Tau = T2 + np.einsum("ai,bj->abij", T1, T1)                                  # :117-118
    FAE, FMI, FME = cc.FAE, cc.FMI, cc.FME
    WAMIESA, WAMEI, WMNEJ, WABEJ, WAMIJ, WMNIJ = cc.WAMIE, cc.WAMEI, cc.WMNEJ, cc.WABEJ, cc.WAMIJ, cc.WMNIJ
    es = np.einsum
    R2SA = 2 * R2 - R2.transpose(0, 1, 3, 2)                                     # :381-382
    XMI = es("efmn,efin->mi", VABIJ, R2SA)                                       # :384
    Z1 = (es("ae,ei->ai", FAE, R1) - es("mi,am->ai", FMI, R1) - es("mi,am->ai", XMI, T1)
          + es("amie,em->ai", WAMIESA, R1) + es("me,aeim->ai", FME, R2SA)
          + es("efam,efim->ai", VABCI, R2SA) - es("mnei,eamn->ai", WMNEJ, R2SA))   # :386-392
    XMI = XMI + 2 * es("nmei,en->mi", WMNEJ, R1) - es("mnei,en->mi", WMNEJ, R1)  # :394-395
    XME = 2 * es("efmn,fn->me", VABIJ, R1) - es("efnm,fn->me", VABIJ, R1)        # :397-398
    XAE = (2 * es("efam,fm->ae", VABCI, R1) - es("feam,fm->ae", VABCI, R1)
           - es("me,am->ae", XME, T1) - es("efmn,afmn->ae", VABIJ, R2SA))        # :400-403
    XAMIJ = es("efam,efij->amij", VABCI, R2)                                     # :405
    XMNIJ = es("efmn,efij->mnij", VABIJ, R2)                                     # :407
    Z2 = (es("abej,ei->abij", WABEJ, R1) - es("amij,bm->abij", WAMIJ, R1)
          - es("amij,bm->abij", XAMIJ, T1) + es("ae,ebij->abij", FAE, R2)
          - es("mi,abmj->abij", FMI, R2) + es("ae,ebij->abij", XAE, T2)
          - es("mi,abmj->abij", XMI, T2) + 0.5 * es("mnij,abmn->abij", WMNIJ, R2)
          + 0.5 * es("mnij,abmn->abij", XMNIJ, Tau) + 0.5 * es("abef,efij->abij", VABCD, R2)
          + 0.5 * es("amie,ebmj->abij", WAMIESA, R2SA) - 0.5 * es("amei,ebjm->abij", WAMEI, R2)
          - es("amej,ebim->abij", WAMEI, R2))                                    # :409-421
    Z2 = Z2 + Z2.transpose(1, 0, 3, 2)                                           # :423
    return Z1, Z2
