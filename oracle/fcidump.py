"""FCIDUMP reading and the MO-basis closed-shell Fock build, restated on
numpy (oracle; test infrastructure only).

Follows the reference's FCIDUMP task, src/operator/fcidump.cxx:43-517:
header NORB / NELEC; lines "value p q r s" (1-based) with p = 0 the core
energy, q = 0 ignored, r = 0 one-electron integrals (mirrored unless the file
holds both triangles), else (pq|rs) expanded to its eight symmetric positions,
lines whose ordered pair (p,q) precedes (r,s) skipped.  The Fock matrix of
the first nocc doubly occupied orbitals is f = h + 2J - K and the SCF energy
E = E_core + sum_i (h_ii + f_ii) (the reference accumulates the same sums line
by line, fcidump.cxx:125-300).

Pinned by tests/test_oracle_fcidump.py: a file written from converged SCF
orbitals of H2O/DZ (bohr geometry) reproduces the reference's SCF energy
and MP2 (test/input.aq:42-43); a hand-written file with one integral of each
kind; the eight-fold expansion.
"""

import re

import numpy as np


def read(path, full_1e=False):
    text = open(path).read().splitlines()
    header = []
    k = 0
    for k, line in enumerate(text):
        header.append(line)
        if re.search(r"(/|[$&]END)", line, re.I):
            break
    hdr = " ".join(header)
    norb = int(re.search(r"NORB\s*=\s*([0-9]+)", hdr, re.I).group(1))
    nelec = int(re.search(r"NELEC\s*=\s*([0-9]+)", hdr, re.I).group(1))
    h = np.zeros((norb, norb))
    eri = np.zeros((norb,) * 4)
    ecore = 0.0
    for line in text[k + 1:]:
        tok = line.split()
        if not tok:
            continue
        val = float(tok[0].replace("D", "E").replace("d", "e"))
        p, q, r, s = (int(t) for t in tok[1:5])
        if p == 0:
            ecore += val
        elif q == 0:
            continue
        elif r == 0:
            h[p - 1, q - 1] += val
            if p != q and not full_1e:
                h[q - 1, p - 1] += val
        else:
            p, q = max(p, q), min(p, q)
            r, s = max(r, s), min(r, s)
            if (p, q) < (r, s):
                continue
            a, b, c, d = p - 1, q - 1, r - 1, s - 1
            for x in ((a, b, c, d), (b, a, c, d), (a, b, d, c), (b, a, d, c),
                      (c, d, a, b), (d, c, a, b), (c, d, b, a), (d, c, b, a)):
                eri[x] = val
    return h, eri, ecore, norb, nelec


def closed_shell(h, eri, nocc):
    """(f, E_electronic) for the first nocc orbitals doubly occupied."""
    o = slice(0, nocc)
    J = np.einsum("pqkk->pq", eri[:, :, o, o])
    K = np.einsum("pkkq->pq", eri[:, o, o, :])
    f = h + 2 * J - K
    e = float(np.trace(h[o, o]) + np.trace(f[o, o]))
    return f, e


def blocks(h, eri, nocc):
    """Fock blocks + seven Dirac blocks in the layout oracle.rhfccsd takes."""
    from .synth import blocks_from_eri
    f, e = closed_shell(h, eri, nocc)
    o = nocc
    fock = {"fAI": f[o:, :o].copy(), "fAB": f[o:, o:].copy(), "fIJ": f[:o, :o].copy()}
    return fock, blocks_from_eri(eri, o), e
