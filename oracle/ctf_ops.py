"""CTF operator semantics on numpy arrays (oracle; test infrastructure only).

A tensor is a numpy float64 array whose axis k carries the k-th index label.
The engine stores the same tensor column-major (first index fastest), i.e. the
CTF key convention key = sum_k idx_k * prod_{j<k} len_j
(reference: src/tensor/ctf_tensor.cxx:493-501, ctf_tensor.hpp:219-228);
``to_keys``/``from_flat`` convert between the two.

Every function here states the reference routine it restates.  Label rules
follow CTF's einsum semantics as used through ``CTFTensor``:

* a label present in both inputs and absent from the output is contracted;
* a label present in one input only and absent from the output is summed;
* a label present in the output only is broadcast;
* a label repeated inside one operand addresses that operand's diagonal.
"""

import numpy as np
from numpy.lib.stride_tricks import as_strided

__all__ = ["label_view", "contract", "tsum", "scale", "fill", "dot", "weight",
           "norm", "div", "invert", "to_flat", "from_flat", "read_keys", "write_keys"]


def label_view(arr, labels):
    """Return (view, unique_labels): a strided view of ``arr`` with one axis per
    distinct label.  Repeated labels collapse onto the diagonal (stride = sum of
    the strides), which is how CTF addresses e.g. ``A["ii"]``.  The view is
    writeable when ``arr`` is."""
    if len(labels) != arr.ndim:
        raise ValueError(f"index string '{labels}' has {len(labels)} labels, tensor has {arr.ndim} dims")
    uniq, shape, strides = [], [], []
    for k, c in enumerate(labels):
        if c in uniq:
            j = uniq.index(c)
            if shape[j] != arr.shape[k]:
                raise ValueError(f"label '{c}' has inconsistent lengths")
            strides[j] += arr.strides[k]
        else:
            uniq.append(c)
            shape.append(arr.shape[k])
            strides.append(arr.strides[k])
    return as_strided(arr, shape=shape, strides=strides, writeable=arr.flags.writeable), "".join(uniq)


def _reduce_to(arr, labels, keep):
    """Sum ``arr`` (unique labels) over every label not in ``keep``."""
    drop = tuple(k for k, c in enumerate(labels) if c not in keep)
    if drop:
        arr = arr.sum(axis=drop)
    return arr, "".join(c for c in labels if c in keep)


def _assign(beta, C, idxC, alpha, R, idxR):
    """C[idxC] = beta*C + alpha*R[idxR], R broadcast over C-only labels.
    R is fully evaluated before C is touched (snapshot semantics)."""
    Cv, cl = label_view(C, idxC)
    # bring R to C's label order, inserting unit axes for broadcast labels
    order = [idxR.index(c) for c in cl if c in idxR]
    R = np.transpose(R, order) if order else R
    shape = [Cv.shape[k] if c not in idxR else -1 for k, c in enumerate(cl)]
    if any(s != -1 for s in shape):
        rs = []
        it = iter(R.shape)
        for s in shape:
            rs.append(next(it) if s == -1 else 1)
        R = R.reshape(rs)
    R = np.array(R, copy=True)
    if beta == 0.0:
        Cv[...] = alpha * R
    else:
        Cv *= beta
        Cv += alpha * R


def contract(alpha, A, idxA, B, idxB, beta, C, idxC):
    """C[idxC] = beta*C[idxC] + alpha * A[idxA] * B[idxB].

    Restates CTFTensor::mult (src/tensor/ctf_tensor.cxx:434-442), reached from
    ``C[..] (=|+=|-=) k*A[..]*B[..]`` via indexable_tensor.hpp:321-346
    (``=`` gives beta=0, ``+=`` beta=1, ``-=`` negates alpha)."""
    Av, la = label_view(A, idxA)
    Bv, lb = label_view(B, idxB)
    _, lc = label_view(C, idxC)
    # labels in one input only and not in C are summed first
    Av, la = _reduce_to(Av, la, set(lb) | set(lc))
    Bv, lb = _reduce_to(Bv, lb, set(la) | set(lc))
    out = "".join(c for c in lc if c in la or c in lb)
    R = np.einsum(f"{la},{lb}->{out}", Av, Bv, optimize=True)
    _assign(beta, C, idxC, alpha, R, out)


def tsum(alpha, A, idxA, beta, B, idxB):
    """B[idxB] = beta*B[idxB] + alpha*A[idxA] (permuted axpy).

    Restates CTFTensor::sum (ctf_tensor.cxx:454-458).  Correct when A is B
    (the aliased ``Z2["abij"] += Z2["baji"]``, rhfccsd.cxx:262,309): the
    permuted source is snapshotted before B is written."""
    Av, la = label_view(A, idxA)
    _, lb = label_view(B, idxB)
    Av, la = _reduce_to(Av, la, set(lb))
    _assign(beta, B, idxB, alpha, np.array(Av, copy=True), la)


def scale(alpha, X, idx):
    """X[idx] = alpha*X[idx] (ctf_tensor.cxx:461-464); a repeated label scales
    only the diagonal it addresses."""
    Xv, _ = label_view(X, idx)
    Xv *= alpha


def fill(X, alpha, beta):
    """X = beta*X + alpha (scalar broadcast sum, ctf_tensor.cxx:445-451)."""
    X *= beta
    X += alpha


def dot(A, idxA, B, idxB):
    """sum over all labels of A[idxA]*B[idxB] (ctf_tensor.cxx:467-478; the
    ``scalar(A*B)`` front end is indexable_tensor.hpp:514-518)."""
    Av, la = label_view(A, idxA)
    Bv, lb = label_view(B, idxB)
    return float(np.einsum(f"{la},{lb}->", Av, Bv, optimize=True))


def weight(X, d, shift=0.0):
    """Divide each element by (sum_k d_k[i_k] + shift); an element whose
    |denominator| < 1e-4 is set to zero.  Restates CTFTensor::weight
    (ctf_tensor.cxx:481-514)."""
    if X.ndim == 0:
        return
    if len(d) != X.ndim:
        raise ValueError("need one denominator vector per dimension")
    den = np.zeros(X.shape)
    for k, dk in enumerate(d):
        dk = np.asarray(dk, dtype=np.float64)
        if dk.shape != (X.shape[k],):
            raise ValueError("denominator length mismatch")
        sh = [1] * X.ndim
        sh[k] = X.shape[k]
        den = den + dk.reshape(sh)
    den = den + shift
    small = np.abs(den) < 1e-4
    safe = np.where(small, 1.0, den)
    X[...] = np.where(small, 0.0, X / safe)


def norm(X, p):
    """p=0 (written ``00`` in the reference): max|x|; p=1: sum|x|; p=2: 2-norm
    (CTFTensor::norm, ctf_tensor.cxx:403-431)."""
    if p == 0:
        return float(np.max(np.abs(X))) if X.size else 0.0
    if p == 1:
        return float(np.sum(np.abs(X)))
    if p == 2:
        return float(np.sqrt(np.sum(X * X)))
    raise ValueError("p must be 0, 1 or 2")


DBL_MIN = np.finfo(np.float64).tiny


def div(alpha, A, B, beta, C):
    """C = beta*C + alpha*A/B elementwise, skipping |B| <= DBL_MIN
    (CTFTensor::div, ctf_tensor.cxx:271-342)."""
    m = np.abs(B) > DBL_MIN
    C[m] = beta * C[m] + alpha * A[m] / B[m]


def invert(alpha, A, beta, C):
    """C = beta*C + alpha/A elementwise, skipping |A| <= DBL_MIN
    (CTFTensor::invert, ctf_tensor.cxx:344-371)."""
    m = np.abs(A) > DBL_MIN
    C[m] = beta * C[m] + alpha / A[m]


def to_flat(X):
    """Column-major flat copy (CTF key order)."""
    return np.asarray(X, dtype=np.float64).ravel(order="F").copy()


def from_flat(flat, shape):
    return np.asarray(flat, dtype=np.float64).reshape(shape, order="F").copy()


def read_keys(X, keys):
    """Values at column-major keys (CTFTensor::getRemoteData, ctf_tensor.hpp:139-160)."""
    return X.ravel(order="F")[np.asarray(keys, dtype=np.int64)].copy()


def write_keys(X, keys, vals, alpha=1.0, beta=0.0):
    """x[key] = beta*x[key] + alpha*val (CTF write with alpha/beta,
    ctf_tensor.hpp:162-180)."""
    idx = np.unravel_index(np.asarray(keys, dtype=np.int64), X.shape, order="F")
    X[idx] = beta * X[idx] + alpha * np.asarray(vals, dtype=np.float64)
