"""Sum-of-products Hamiltonian -> MPO by successive QR decompositions of the
coefficient matrix between left and right operator strings (reference
mps/symbolic_mpo.py: construct_symbolic_mpo 22-161, _decompose_qr 291-347,
_terms_to_table 356-417, _deduplicate_table 420-438, compose_symbolic_mo
443-455, symbolic_mo_to_numeric_mo 457-468).

The same steps are kept so that the bond dimensions (ranks of the
coefficient matrices at each cut, rtol 1e-10) and the operator structure
match the reference: rows of the term table are deduplicated and sorted
(numpy.unique), unique left strings are sorted, unique right strings keep
first-appearance order, and the rank-revealing step is scipy's pivoted QR.
Host-only float64 code, run once per model.
"""
from collections import namedtuple
from typing import List

import numpy as np
import scipy.linalg
import scipy.sparse

from renormalizer_amd.model.op import Op

OpTuple = namedtuple("OpTuple", ["symbol", "qn", "factor"])

_RTOL = 1e-10
_ATOL = 1e-10


def terms_to_table(model, terms: List[Op], const: float):
    """Table of primary-operator indices (one row per term, one column per
    site) + factors.  Index i < nsite is the identity of site i."""
    primary_each_site = []
    primary_ops = []
    identity_row = []
    for idx, b in enumerate(model.basis):
        dof = b.dof[0] if b.multi_dof else b.dof
        op = Op.identity(dof, qn_size=model.qn_size)
        primary_each_site.append({op: idx})
        primary_ops.append(op)
        identity_row.append(idx)
    table, factors = [], []
    for op in terms:
        elem_ops, factor = op.split_elementary(model.dof_to_siteidx)
        row = identity_row.copy()
        for e in elem_ops:
            site = model.dof_to_siteidx[e.dofs[0]]
            if e not in primary_each_site[site]:
                primary_each_site[site][e] = len(primary_ops)
                primary_ops.append(e)
            row[site] = primary_each_site[site][e]
        table.append(row)
        factors.append(factor)
    if const != 0:
        table.append(identity_row.copy())
        factors.append(const)
    assert len(primary_ops) < np.iinfo(np.uint16).max
    table = np.array(table, dtype=np.uint16)
    factor = np.array(factors)
    table, factor = deduplicate_table(table, factor)
    return table, primary_ops, factor


def deduplicate_table(table, factor):
    """Sum the factors of identical rows (rows come out sorted) and drop
    terms below 1e-15 of the largest factor."""
    new_table, inverse = np.unique(table, axis=0, return_inverse=True)
    inverse = np.asarray(inverse).reshape(-1)
    summed = np.zeros(len(new_table), dtype=np.result_type(factor.dtype, np.float64))
    np.add.at(summed, inverse, factor)
    keep = np.abs(summed) > np.max(np.abs(summed)) * 1e-15
    return new_table[keep], summed[keep]


def _compute_qn(in_ops, symbol, primary_ops):
    return in_ops[symbol[0]][0].qn + primary_ops[symbol[1]].qn


def _decompose_site(table, in_ops, factor, primary_ops):
    """One cut: split each row into (left string: [in_op, this site's op],
    right string: rest), QR the coefficient matrix Gamma[left, right]."""
    table_row = table[:, :2]
    table_col = table[:, 2:]
    term_row, row_inv = np.unique(table_row, axis=0, return_inverse=True)
    row_inv = np.asarray(row_inv).reshape(-1)
    col_index = {}
    term_col = []
    col_inv = np.empty(len(table_col), dtype=np.int64)
    for i, row in enumerate(table_col):
        key = row.tobytes()
        j = col_index.get(key)
        if j is None:
            j = len(term_col)
            col_index[key] = j
            term_col.append(row)
        col_inv[i] = j
    gamma = np.zeros((len(term_row), len(term_col)), dtype=factor.dtype)
    gamma[row_inv, col_inv] = factor  # rows of the table are unique: no collisions
    if gamma.shape[1] != 1:
        q, r, p = scipy.linalg.qr(gamma, mode="economic", pivoting=True)
    else:
        q, r, p = gamma, np.array([[1.0]]), np.array([0])
    rank = int(np.sum(np.abs(np.diag(r)) > np.abs(r[0][0]) * _RTOL))
    out_ops = [[] for _ in range(rank)]
    for i, j in zip(*np.where(np.abs(q[:, :rank]) > _ATOL)):
        symbol = term_row[i]
        out_ops[j].append(OpTuple(symbol, _compute_qn(in_ops, symbol, primary_ops), q[i, j]))
    r2 = r[:rank, np.argsort(p)]
    idx1, idx2 = np.where(np.abs(r2) > np.abs(r[0][0]) * _RTOL)
    new_factor = r2[(idx1, idx2)]
    new_table = np.concatenate([idx1.reshape(-1, 1), np.array([term_col[i] for i in idx2]).reshape(len(idx2), -1)],
                               axis=1).astype(np.int64)
    return out_ops, new_table, new_factor


def construct_symbolic_mpo(table, primary_ops, factor):
    """Returns (symbolic mpo, mpo qn list, qntot, qnidx, out_ops_list)."""
    qn_size = len(primary_ops[0].qn)
    nsite = table.shape[1]
    if table.shape[0] == 1:
        mpo = []
        mpoqn = [np.zeros((1, qn_size), dtype=int)]
        qn = np.zeros(qn_size, dtype=int)
        for idx in table[0]:
            op = primary_ops[idx]
            mo = np.full((1, 1), None)
            mo[0][0] = [op]
            mpo.append(mo)
            qn = mpoqn[-1][0] + op.qn
            mpoqn.append(np.array([qn]))
        mpo[-1][0][0][0] = factor[0] * mpo[-1][0][0][0]
        mpoqn[-1] = np.zeros((1, qn_size), dtype=int)
        return mpo, mpoqn, qn, nsite - 1, None
    pad = np.zeros((table.shape[0], 1), dtype=np.int64)
    table = np.concatenate((pad, table.astype(np.int64), pad), axis=1)
    assert len(np.unique(table, axis=0)) == len(table)
    in_ops = [[OpTuple([0], np.zeros(qn_size, dtype=int), 1)]]
    out_ops_list = [in_ops]
    for _ in range(nsite):
        out_ops, table, factor = _decompose_site(table, in_ops, factor, primary_ops)
        in_ops = out_ops
        out_ops_list.append(out_ops)
    assert len(factor) == 1 and len(table) == 1 and factor[0] == 1
    mpo = [compose_symbolic_mo(out_ops_list[i], out_ops_list[i + 1], primary_ops) for i in range(nsite)]
    mpoqn = [np.array([o[0].qn for o in ops]) for ops in out_ops_list]
    qntot = mpoqn[-1][0]
    mpoqn[-1] = np.zeros((1, qn_size), dtype=int)
    return mpo, mpoqn, qntot, nsite - 1, out_ops_list


def compose_symbolic_mo(in_ops, out_ops, primary_ops):
    mo = np.empty((len(in_ops), len(out_ops)), dtype=object)
    for idx in np.ndindex(mo.shape):
        mo[idx] = []
    for iop, out_op in enumerate(out_ops):
        for composed in out_op:
            mo[composed.symbol[0]][iop].append(composed.factor * primary_ops[composed.symbol[1]])
    return mo


def symbolic_mo_to_numeric_mo(basis, mo, dtype):
    """(D_l, D_r) grid of operator sums -> W[D_l, d_out, d_in, D_r]."""
    pdim = basis.nbas
    mat = np.zeros(list(mo.shape) + [pdim, pdim], dtype=dtype)
    for idx in np.ndindex(mo.shape):
        for term in mo[idx]:
            mat[idx] += basis.op_mat(term)
    return mat.transpose(0, 2, 3, 1)
