"""Gaussian-basis integrals, closed-shell SCF and the AO->MO block builder,
restated on numpy (oracle; test infrastructure only -- nothing in the product
imports this module).

What it follows in the reference:
  * GENBAS reader                src/input/basis.cxx:15-75 (element symbol before
                                 ':', comment line, nshell, L[], ncontr[],
                                 nprim[], then exponents and a primitive-major
                                 coefficient table per shell)
  * contraction normalisation    src/integrals/shell.cxx:114-134
  * spherical d/f functions      src/integrals/shell.cxx:139-160 (pure harmonics,
                                 contaminants dropped; spherical is the default,
                                 src/input/molecule.cxx:1364-1367)
  * units                        src/input/molecule.cxx:180-182, 1335-1338
  * 1e/2e integrals              src/integrals/1eints.cxx, os.cxx (the values,
                                 not the recursion: here McMurchie-Davidson
                                 Hermite expansions, Helgaker-Joergensen-Olsen,
                                 "Molecular Electronic-Structure Theory" ch. 9)
  * SCF                          src/scf/uhf.cxx:17-409 (closed shell: UHF with
                                 equal alpha/beta = RHF), convergence 1e-12,
                                 frozen_core false (src/scf/aouhf.cxx:266-282)
  * AO->MO                       src/operator/rhfaomoints.cxx:38-180: the seven
                                 Dirac blocks <pq|rs> = (pr|qs) and the Fock
                                 blocks f(AI), f(AB), f(IJ) in the MO basis.

Energies are invariant to the AO ordering and to the normalisation of each AO
function, so neither is taken from the reference.  Pinned by the reference's
known answers (test/input.aq:20-23, 42-43): SCF, MP2 and CCSD/CCD correlation
energies of H2O/cc-pVDZ and SCF and MP2 of H2O/DZ at the bohr geometry, each to
1e-9 Eh; the Boys function against direct quadrature; overlap and kinetic
matrices against closed forms.
"""

import math
import re

import numpy as np
from scipy import special

ANGSTROM_TO_BOHR = 1.88972612456506198632428439     # molecule.cxx:1335-1336

SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar "
           "K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr").split()


def atomic_number(symbol):
    return [s.upper() for s in SYMBOLS].index(symbol.upper()) + 1


# ---------------------------------------------------------------- GENBAS ----

def parse_genbas(text):
    """{ELEMENT: [(L, exps[nprim], coefs[nprim, ncontr]), ...]} (basis.cxx:15-75).

    The coefficient table is stored primitive-major (one row per primitive,
    one column per contraction), as in the file."""
    lines = text.splitlines()
    out = {}
    i = 0
    while i < len(lines):
        line = lines[i]
        i += 1
        if not line.strip() or line[0] == "!":
            continue
        sep = line.find(":")
        if sep < 1 or sep > 2:
            raise ValueError(f"GENBAS line {i}: element header expected")
        elem = line[:sep].strip().upper()
        i += 1                                   # comment line
        numbers = []

        def take(n):
            nonlocal i
            while len(numbers) < n:
                if i >= len(lines):
                    raise ValueError("GENBAS: premature end of file")
                numbers.extend(lines[i].split())
                i += 1
            vals = numbers[:n]
            del numbers[:n]
            return vals

        nshell = int(take(1)[0])
        Ls = [int(x) for x in take(nshell)]
        ncontr = [int(x) for x in take(nshell)]
        nprim = [int(x) for x in take(nshell)]
        shells = []
        for L, nc, npr in zip(Ls, ncontr, nprim):
            exps = np.array([float(x) for x in take(npr)])
            coef = np.array([float(x) for x in take(npr * nc)]).reshape(npr, nc)
            shells.append((L, exps, coef))
        assert not numbers
        out[elem] = shells
    return out


# ------------------------------------------------------- angular parts ------

def cart_powers(L):
    """Cartesian components of a shell: x descending, then y descending."""
    return [(x, y, L - x - y) for x in range(L, -1, -1) for y in range(L - x, -1, -1)]


def dfact(n):
    return 1.0 if n <= 0 else float(np.prod(np.arange(n, 0, -2, dtype=np.float64)))


def cart_overlap_axial(L):
    """<c|c'> of unit-normalised-axial cartesian Gaussians of one exponent and
    centre: prod_d (p_d+p'_d-1)!! / (2L-1)!! when every p_d+p'_d is even."""
    P = cart_powers(L)
    S = np.zeros((len(P), len(P)))
    for a, pa in enumerate(P):
        for b, pb in enumerate(P):
            s = [pa[d] + pb[d] for d in range(3)]
            if all(x % 2 == 0 for x in s):
                S[a, b] = dfact(s[0] - 1) * dfact(s[1] - 1) * dfact(s[2] - 1) / dfact(2 * L - 1)
    return S


def solid_harmonics(L):
    """Rows m = -L..L: coefficients of the real solid harmonic S_Lm over the
    cartesian monomials of cart_powers(L) (Helgaker et al. eq. 6.4.47-6.4.48),
    renormalised to unit norm over axially normalised cartesian Gaussians."""
    P = {p: k for k, p in enumerate(cart_powers(L))}
    if L < 2:
        # s: 1; p: y, z, x in m order -1, 0, 1 (any basis of the shell works)
        if L == 0:
            return np.ones((1, 1))
        M = np.zeros((3, 3))
        M[0, P[(0, 1, 0)]] = M[1, P[(0, 0, 1)]] = M[2, P[(1, 0, 0)]] = 1.0
        return M
    M = np.zeros((2 * L + 1, len(P)))
    for m in range(-L, L + 1):
        am = abs(m)
        vm = 0.5 if m < 0 else 0.0
        row = M[m + L]
        for t in range((L - am) // 2 + 1):
            for u in range(t + 1):
                v = vm
                while v <= math.floor(am / 2 - vm) + vm + 1e-9:
                    c = ((-1) ** int(round(t + v - vm)) * 0.25 ** t * special.comb(L, t, exact=True)
                         * special.comb(L - t, am + t, exact=True) * special.comb(t, u, exact=True)
                         * special.comb(am, int(round(2 * v)), exact=True))
                    px = 2 * t + am - int(round(2 * (u + v)))
                    py = int(round(2 * (u + v)))
                    pz = L - 2 * t - am
                    row[P[(px, py, pz)]] += c
                    v += 1.0
    S = cart_overlap_axial(L)
    norms = np.sqrt(np.einsum("mi,ij,mj->m", M, S, M))
    return M / norms[:, None]


# ------------------------------------------------------------- basis --------

class Basis:
    """Shells of a molecule, the primitive cartesian Gaussians they are built
    from, and the matrix mapping primitives to (spherical, normalised) AOs."""

    def __init__(self, atoms, basis_sets, spherical=True):
        """atoms: [(Z, (x, y, z) in bohr)]; basis_sets: {ELEMENT: shells}."""
        self.atoms = [(int(Z), np.asarray(r, dtype=np.float64)) for Z, r in atoms]
        self.shells = []                       # (atom, L, exps, coefs[nprim, ncontr])
        for ia, (Z, r) in enumerate(self.atoms):
            for L, exps, coef in basis_sets[SYMBOLS[Z - 1].upper()]:
                self.shells.append((ia, L, exps, self._normalise(L, exps, coef)))
        prim_center, prim_exp, prim_pow, rows, cols, vals = [], [], [], [], [], []
        nao = 0
        for ia, L, exps, coef in self.shells:
            T = solid_harmonics(L) if spherical else np.eye(len(cart_powers(L)))
            nf = T.shape[0]
            base = len(prim_exp)
            powers = cart_powers(L)
            for j, a in enumerate(exps):
                nrm = (2 * a / np.pi) ** 0.75 * (4 * a) ** (L / 2) / math.sqrt(dfact(2 * L - 1))
                for c, pw in enumerate(powers):
                    prim_center.append(self.atoms[ia][1])
                    prim_exp.append(a)
                    prim_pow.append(pw)
                    for k in range(coef.shape[1]):
                        for f in range(nf):
                            rows.append(base + j * len(powers) + c)
                            cols.append(nao + k * nf + f)
                            vals.append(coef[j, k] * nrm * T[f, c])
            nao += coef.shape[1] * nf
        self.nao = nao
        self.center = np.array(prim_center)
        self.exp = np.array(prim_exp)
        self.pow = np.array(prim_pow, dtype=np.int64)
        self.C = np.zeros((len(prim_exp), nao))
        np.add.at(self.C, (np.array(rows), np.array(cols)), np.array(vals))

    @staticmethod
    def _normalise(L, exps, coef):
        """Each contraction to unit norm over normalised primitives
        (the normalisation of shell.cxx:114-134)."""
        a = exps[:, None]
        b = exps[None, :]
        ov = (2 * np.sqrt(a * b) / (a + b)) ** (L + 1.5)
        coef = np.array(coef, dtype=np.float64)
        for k in range(coef.shape[1]):
            coef[:, k] /= math.sqrt(coef[:, k] @ ov @ coef[:, k])
        return coef

    def nuclear_repulsion(self):
        e = 0.0
        for i, (Zi, ri) in enumerate(self.atoms):
            for Zj, rj in self.atoms[:i]:
                e += Zi * Zj / np.linalg.norm(ri - rj)
        return e

    def n_electrons(self, charge=0):
        return sum(Z for Z, _ in self.atoms) - charge


# ------------------------------------------------------ Boys function -------

def boys(nmax, T):
    """F_n(T) = int_0^1 t^(2n) exp(-T t^2) dt for n = 0..nmax; array [nmax+1, *T.shape].
    Series e^{-T} sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1)) below T = 25
    (positive terms, no cancellation), the incomplete-gamma closed form
    Gamma(n+1/2) P(n+1/2, T) / (2 T^(n+1/2)) above."""
    T = np.asarray(T, dtype=np.float64)
    out = np.empty((nmax + 1,) + T.shape)
    small = T < 25.0
    Ts = T[small]
    Tl = T[~small]
    for n in range(nmax + 1):
        if Ts.size:
            term = np.full(Ts.shape, 1.0 / (2 * n + 1))
            acc = term.copy()
            for k in range(1, 160):
                term = term * (2 * Ts) / (2 * n + 2 * k + 1)
                acc += term
            out[n][small] = np.exp(-Ts) * acc
        if Tl.size:
            a = n + 0.5
            out[n][~small] = special.gamma(a) * special.gammainc(a, Tl) / (2 * Tl ** a)
    return out


# ------------------------------------------------ Hermite machinery ---------

def hermite_E(lmax_a, lmax_b, XPA, XPB, p):
    """E[i, j, t] for i <= lmax_a, j <= lmax_b, t <= i + j (Helgaker 9.5.6-9.5.7),
    vectorised over the trailing shape of XPA/XPB/p; E[0,0,0] = 1 (the
    exponential prefactor is applied separately)."""
    shape = np.shape(XPA)
    E = np.zeros((lmax_a + 1, lmax_b + 1, lmax_a + lmax_b + 2) + shape)
    E[0, 0, 0] = 1.0
    inv2p = 0.5 / p
    for i in range(lmax_a + 1):
        for j in range(lmax_b + 1):
            if i == 0 and j == 0:
                continue
            if i > 0:
                src, X, = E[i - 1, j], XPA
            else:
                src, X = E[i, j - 1], XPB
            for t in range(i + j + 1):
                val = X * src[t]
                if t > 0:
                    val = val + inv2p * src[t - 1]
                if t + 1 <= i + j - 1:
                    val = val + (t + 1) * src[t + 1]
                E[i, j, t] = val
    return E


def hermite_R(L, alpha, X, Y, Z):
    """{(t,u,v): R_tuv} for t+u+v <= L (Helgaker 9.9.18-9.9.20), vectorised."""
    T = alpha * (X * X + Y * Y + Z * Z)
    F = boys(L, T)
    prev = None
    for n in range(L, -1, -1):
        cur = {(0, 0, 0): (-2 * alpha) ** n * F[n]}
        top = L - n
        for s in range(1, top + 1):
            for t in range(s, -1, -1):
                for u in range(s - t, -1, -1):
                    v = s - t - u
                    if t > 0:
                        val = X * prev[(t - 1, u, v)]
                        if t > 1:
                            val = val + (t - 1) * prev[(t - 2, u, v)]
                    elif u > 0:
                        val = Y * prev[(t, u - 1, v)]
                        if u > 1:
                            val = val + (u - 1) * prev[(t, u - 2, v)]
                    else:
                        val = Z * prev[(t, u, v - 1)]
                        if v > 1:
                            val = val + (v - 1) * prev[(t, u, v - 2)]
                    cur[(t, u, v)] = val
        prev = cur
    return prev


def _pairs(basis, idx_a, idx_b):
    A = basis.center[idx_a]
    B = basis.center[idx_b]
    a = basis.exp[idx_a]
    b = basis.exp[idx_b]
    p = a + b
    P = (a[:, None] * A + b[:, None] * B) / p[:, None]
    K = np.exp(-(a * b / p) * np.sum((A - B) ** 2, axis=1))
    return A, B, a, b, p, P, K


def _pair_hermite(basis, ia, ib, extra_b=0):
    """Per-pair Hermite tables Ex/Ey/Ez[i, j, t, pair] with i, j up to the
    largest power present (+extra_b on the ket side for kinetic integrals)."""
    A, B, a, b, p, P, K = _pairs(basis, ia, ib)
    lm = int(basis.pow.max())
    Es = [hermite_E(lm, lm + extra_b, P[:, d] - A[:, d], P[:, d] - B[:, d], p) for d in range(3)]
    return Es, p, P, K, b


def _select(E, pa, pb):
    """E[pa[k], pb[k], :, k] -> [t, k]."""
    k = np.arange(len(pa))
    return E[pa, pb, :, k].T


# ----------------------------------------------------- one-electron ---------

def one_electron(basis):
    """(S, T, V) in the AO basis."""
    n = len(basis.exp)
    ia, ib = [x.ravel() for x in np.meshgrid(np.arange(n), np.arange(n), indexing="ij")]
    Es, p, P, K, b = _pair_hermite(basis, ia, ib, extra_b=2)
    pa, pb = basis.pow[ia], basis.pow[ib]
    # 1-D overlaps S_d(i, j) = E^{ij}_0 sqrt(pi/p)
    k = np.arange(len(ia))
    sq = np.sqrt(np.pi / p)

    def s1(d, di, dj):
        j = pb[:, d] + dj
        ok = j >= 0
        val = np.zeros(len(ia))
        val[ok] = Es[d][pa[ok, d], j[ok], 0, k[ok]] * sq[ok]
        return val

    Sd = [s1(d, 0, 0) for d in range(3)]
    Td = []
    for d in range(3):
        j = pb[:, d]
        Td.append(-0.5 * (j * (j - 1) * s1(d, 0, -2) - 2 * b * (2 * j + 1) * Sd[d] + 4 * b * b * s1(d, 0, 2)))
    S = K * Sd[0] * Sd[1] * Sd[2]
    T = K * (Td[0] * Sd[1] * Sd[2] + Sd[0] * Td[1] * Sd[2] + Sd[0] * Sd[1] * Td[2])
    # nuclear attraction
    Ex = _select(Es[0], pa[:, 0], pb[:, 0])
    Ey = _select(Es[1], pa[:, 1], pb[:, 1])
    Ez = _select(Es[2], pa[:, 2], pb[:, 2])
    L = 2 * int(basis.pow.sum(axis=1).max())
    V = np.zeros(len(ia))
    for Zc, C in basis.atoms:
        R = hermite_R(L, p, P[:, 0] - C[0], P[:, 1] - C[1], P[:, 2] - C[2])
        acc = np.zeros(len(ia))
        for (t, u, v), r in R.items():
            if t < Ex.shape[0] and u < Ey.shape[0] and v < Ez.shape[0]:
                acc += Ex[t] * Ey[u] * Ez[v] * r
        V -= Zc * acc
    V *= 2 * np.pi / p * K
    C = basis.C
    tr = lambda M: C.T @ M.reshape(n, n) @ C
    return tr(S), tr(T), tr(V)


# ----------------------------------------------------- two-electron ---------

def two_electron(basis, chunk=64):
    """Full AO tensor eri[m, n, l, s] = (mn|ls) (Mulliken)."""
    n = len(basis.exp)
    iu, ju = np.triu_indices(n)
    npair = len(iu)
    Es, p, P, K, _ = _pair_hermite(basis, iu, ju)
    pa, pb = basis.pow[iu], basis.pow[ju]
    Eh = [_select(Es[d], pa[:, d], pb[:, d]) for d in range(3)]       # [t, pair]
    lmax = int(basis.pow.sum(axis=1).max())
    herm = [(t, u, v) for t in range(2 * lmax + 1) for u in range(2 * lmax + 1 - t)
            for v in range(2 * lmax + 1 - t - u)]
    # Hermite expansion of each pair: Ep[h, pair] = Ex_t Ey_u Ez_v (times K)
    Ep = np.array([Eh[0][t] * Eh[1][u] * Eh[2][v] for t, u, v in herm]) * K
    live = [h for h in range(len(herm)) if np.any(Ep[h] != 0)]
    sign = np.array([(-1) ** (t + u + v) for t, u, v in herm])
    Lt = 4 * lmax
    G = np.zeros((npair, npair))
    for c0 in range(0, npair, chunk):
        c1 = min(npair, c0 + chunk)
        pb_ = p[c0:c1, None]
        pk = p[None, :]
        alpha = pb_ * pk / (pb_ + pk)
        D = P[c0:c1, None, :] - P[None, :, :]
        R = hermite_R(Lt, alpha, D[..., 0], D[..., 1], D[..., 2])
        acc = np.zeros((c1 - c0, npair))
        for h1 in live:
            t, u, v = herm[h1]
            e1 = Ep[h1, c0:c1, None]
            if not np.any(e1):
                continue
            inner = np.zeros((c1 - c0, npair))
            for h2 in live:
                t2, u2, v2 = herm[h2]
                inner += (sign[h2] * Ep[h2])[None, :] * R[(t + t2, u + u2, v + v2)]
            acc += e1 * inner
        G[c0:c1] = acc * 2 * np.pi ** 2.5 / (pb_ * pk * np.sqrt(pb_ + pk))
    Gs = np.zeros((n, n, npair))
    Gs[iu, ju] = G
    Gs[ju, iu] = G
    full = np.zeros((n, n, n, n))
    full[:, :, iu, ju] = Gs
    full[:, :, ju, iu] = Gs
    C = basis.C
    full = np.tensordot(full, C, axes=([3], [0]))          # [m, n, l, S]
    full = np.tensordot(full, C, axes=([2], [0]))          # [m, n, S, L]
    full = full.transpose(0, 1, 3, 2)                      # [m, n, L, S]
    full = np.tensordot(C, full, axes=([0], [1]))          # [N, m, L, S]
    full = np.tensordot(C, full, axes=([0], [1]))          # [M, N, L, S]
    return np.ascontiguousarray(full)


# ------------------------------------------------------------- SCF ----------

def rhf(S, H, eri, nocc, enuc, convergence=1e-12, max_iterations=150, diis_order=6):
    """Closed-shell SCF (the closed-shell case of uhf.cxx:17-409): core guess,
    Pulay DIIS on the commutator FDS - SDF, stop when the energy change and the
    largest commutator element are both below `convergence`.

    Returns (E_total, eps[n], C[nao, n], F_ao)."""
    s, U = np.linalg.eigh(S)
    X = U @ np.diag(s ** -0.5) @ U.T

    def diag(F):
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        return e, X @ Cp

    e, C = diag(H)
    E_old = 0.0
    errs, focks = [], []
    for it in range(max_iterations):
        D = C[:, :nocc] @ C[:, :nocc].T
        J = np.einsum("mnls,ls->mn", eri, D)
        K = np.einsum("mlns,ls->mn", eri, D)
        F = H + 2 * J - K
        E = np.sum(D * (H + F)) + enuc
        err = F @ D @ S - S @ D @ F
        focks.append(F)
        errs.append(err)
        focks, errs = focks[-diis_order:], errs[-diis_order:]
        if abs(E - E_old) < convergence and np.abs(err).max() < convergence * 100:
            return E, e, C, F
        E_old = E
        m = len(focks)
        B = -np.ones((m + 1, m + 1))
        B[m, m] = 0.0
        for i in range(m):
            for j in range(m):
                B[i, j] = np.sum(errs[i] * errs[j])
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        c = np.linalg.lstsq(B, rhs, rcond=None)[0][:m]
        e, C = diag(sum(ci * Fi for ci, Fi in zip(c, focks)))
    raise RuntimeError("SCF did not converge")


# ------------------------------------------------------------ AO -> MO ------

def mo_blocks(eri_ao, C, F_ao, nocc):
    """Fock blocks and the seven Dirac blocks (rhfaomoints.cxx:38-180) in the
    layout oracle.rhfccsd.RHFCCSD takes."""
    from .synth import blocks_from_eri
    eri = np.einsum("mnls,mp->pnls", eri_ao, C, optimize=True)
    eri = np.einsum("pnls,nq->pqls", eri, C, optimize=True)
    eri = np.einsum("pqls,lr->pqrs", eri, C, optimize=True)
    eri = np.einsum("pqrs,st->pqrt", eri, C, optimize=True)
    f = C.T @ F_ao @ C
    o = nocc
    fock = {"fAI": f[o:, :o].copy(), "fAB": f[o:, o:].copy(), "fIJ": f[:o, :o].copy()}
    return fock, blocks_from_eri(eri, o)


def molecule(spec_atoms, genbas_text, units="bohr", spherical=True):
    """Basis for [(symbol, x, y, z)] in `units` ('bohr' or 'angstrom')."""
    scale = ANGSTROM_TO_BOHR if units == "angstrom" else 1.0
    atoms = [(atomic_number(s), (x * scale, y * scale, z * scale)) for s, x, y, z in spec_atoms]
    return Basis(atoms, parse_genbas(genbas_text), spherical)


def uhf(S, H, eri, nalpha, nbeta, enuc, convergence=1e-12, max_iterations=300, diis_order=8):
    """Unrestricted SCF (uhf.cxx:17-409): core guess for both spins, Pulay
    DIIS on the stacked orthonormal-basis commutators, stop when the energy
    change and the largest commutator element are below `convergence`
    (x100 for the commutator).

    Returns (E_total, (eps_a, eps_b), (Ca, Cb), (Fa, Fb))."""
    s, U = np.linalg.eigh(S)
    X = U @ np.diag(s ** -0.5) @ U.T

    def diag(F):
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        return e, X @ Cp

    ea, Ca = diag(H)
    eb, Cb = ea.copy(), Ca.copy()
    E_old = 0.0
    hist_f, hist_e = [], []
    for it in range(max_iterations):
        Da = Ca[:, :nalpha] @ Ca[:, :nalpha].T
        Db = Cb[:, :nbeta] @ Cb[:, :nbeta].T
        J = np.einsum("mnls,ls->mn", eri, Da + Db)
        Ka = np.einsum("mlns,ls->mn", eri, Da)
        Kb = np.einsum("mlns,ls->mn", eri, Db)
        Fa = H + J - Ka
        Fb = H + J - Kb
        E = 0.5 * np.sum(Da * (H + Fa)) + 0.5 * np.sum(Db * (H + Fb)) + enuc
        err_a = X.T @ (Fa @ Da @ S - S @ Da @ Fa) @ X
        err_b = X.T @ (Fb @ Db @ S - S @ Db @ Fb) @ X
        conv = max(np.abs(err_a).max(), np.abs(err_b).max())
        if it > 0 and abs(E - E_old) < convergence and conv < 100 * convergence:
            ea, Ca = diag(Fa)
            eb, Cb = diag(Fb)
            return E, (ea, eb), (Ca, Cb), (Fa, Fb)
        E_old = E
        hist_f.append((Fa, Fb))
        hist_e.append(np.concatenate([err_a.ravel(), err_b.ravel()]))
        hist_f, hist_e = hist_f[-diis_order:], hist_e[-diis_order:]
        m = len(hist_f)
        if m >= 2:
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0.0
            for i in range(m):
                for j in range(m):
                    B[i, j] = hist_e[i] @ hist_e[j]
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            c = np.linalg.lstsq(B, rhs, rcond=None)[0][:m]
            Fa = sum(ci * f[0] for ci, f in zip(c, hist_f))
            Fb = sum(ci * f[1] for ci, f in zip(c, hist_f))
        ea, Ca = diag(Fa)
        eb, Cb = diag(Fb)
    raise RuntimeError("UHF did not converge")
