"""Independent determinant-basis FCI used to pin the oracle (tests only).

H = sum_pq T_pq E_pq + 1/2 sum_pqrs V_prqs (E_pq E_rs - delta_qr E_ps) with spin-summed
E_pq and physicist V_ijkl = <ij|kl>. The lowest state of total spin S is taken in the
S_z = S sector with the penalty lambda * S_+^T S_+, whose kernel there is exactly spin S.
Abelian point-group symmetry: the product (XOR) of the irreps of all occupied
spin-orbitals must equal the target irrep.
"""
import itertools

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _strings(L, n):
    return [sum(1 << i for i in c) for c in itertools.combinations(range(L), n)]


def _irrep(s, irreps):
    r = 0
    i = 0
    while s:
        if s & 1:
            r ^= irreps[i]
        s >>= 1
        i += 1
    return r


def _sign(s, p):
    """(-1)^(number of occupied orbitals below p)."""
    return -1 if bin(s & ((1 << p) - 1)).count("1") % 2 else 1


def _basis(L, na, nb, irreps, target):
    A = _strings(L, na)
    B = _strings(L, nb)
    dets = []
    for a in A:
        ia = _irrep(a, irreps) if irreps is not None else 0
        for b in B:
            if irreps is None or ia ^ _irrep(b, irreps) == target:
                dets.append((a, b))
    return dets, {d: i for i, d in enumerate(dets)}


def _epq(dets, index, p, q, na, nb):
    """Spin-summed E_pq = a+_pa a_qa + a+_pb a_qb as a sparse matrix (with the
    convention that the alpha string precedes the beta string)."""
    rows, cols, vals = [], [], []
    for j, (a, b) in enumerate(dets):
        for spin in (0, 1):
            s = a if spin == 0 else b
            if not (s >> q) & 1:
                continue
            sg = _sign(s, q)
            s2 = s ^ (1 << q)
            if (s2 >> p) & 1:
                continue
            sg *= _sign(s2, p)
            s2 |= 1 << p
            d2 = (s2, b) if spin == 0 else (a, s2)
            i = index.get(d2)
            if i is None:
                continue
            rows.append(i)
            cols.append(j)
            vals.append(float(sg))
    n = len(dets)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def fci_energy(L, irreps, N, TwoS, Irrep, T, V, econst=0.0, penalty=4.0, k=4):
    """Lowest eigenvalue of total spin TwoS/2, N electrons and irrep Irrep."""
    T = np.asarray(T, dtype=np.float64).reshape(L, L)
    V = np.asarray(V, dtype=np.float64).reshape(L, L, L, L)
    na = (N + TwoS) // 2
    nb = (N - TwoS) // 2
    # The products E_pq E_rs pass through other irreps, so H is built in the whole
    # S_z sector and restricted to the target irrep at the end.
    dets, index = _basis(L, na, nb, None, None)
    n = len(dets)
    keep = np.array([_irrep(a, irreps) ^ _irrep(b, irreps) == Irrep for (a, b) in dets])
    if not keep.any():
        raise ValueError("empty FCI space")
    E = [[None] * L for _ in range(L)]
    for p in range(L):
        for q in range(L):
            E[p][q] = _epq(dets, index, p, q, na, nb)
    H = sp.csr_matrix((n, n))
    for p in range(L):
        for q in range(L):
            if T[p, q] != 0.0:
                H = H + T[p, q] * E[p][q]
    # (pq|rs) = V[p, r, q, s]
    for p in range(L):
        for q in range(L):
            acc = sp.csr_matrix((n, n))
            for r in range(L):
                for s in range(L):
                    v = V[p, r, q, s]
                    if v != 0.0:
                        acc = acc + v * E[r][s]
            if acc.nnz:
                H = H + 0.5 * (E[p][q] @ acc)
    for p in range(L):
        for s in range(L):
            c = sum(V[p, q, q, s] for q in range(L))
            if c != 0.0:
                H = H - 0.5 * c * E[p][s]
    # S_+ = sum_p a+_pa a_pb into the (na+1, nb-1) sector
    if nb > 0 and na + 1 <= L:
        dets2, index2 = _basis(L, na + 1, nb - 1, None, None)
        rows, cols, vals = [], [], []
        for j, (a, b) in enumerate(dets):
            for p in range(L):
                if not (b >> p) & 1 or (a >> p) & 1:
                    continue
                # a+_pa a_pb |a b>: a_pb passes the whole alpha string (a constant sign,
                # irrelevant inside S_+^T S_+) and the beta electrons below p; a+_pa
                # passes the alpha electrons below p.
                sg = _sign(b, p) * _sign(a, p)
                b2 = b ^ (1 << p)
                a2 = a | (1 << p)
                i = index2.get((a2, b2))
                if i is None:
                    continue
                rows.append(i)
                cols.append(j)
                vals.append(float(sg))
        Sp = sp.csr_matrix((vals, (rows, cols)), shape=(len(dets2), n))
        H = H + penalty * (Sp.T @ Sp)
    H = 0.5 * (H + H.T)
    sel = np.nonzero(keep)[0]
    H = H.tocsr()[sel][:, sel]
    n = len(sel)
    if n <= 600:
        w = np.linalg.eigvalsh(H.toarray())
        return float(w[0]) + econst
    w = spla.eigsh(H, k=min(k, n - 1), which="SA", tol=1e-12, return_eigenvectors=False)
    return float(np.min(w)) + econst


def hubbard(L, t=1.0, U=4.0, periodic=False):
    """Open (or periodic) 1D Hubbard chain in the physicist convention."""
    T = np.zeros((L, L))
    for i in range(L - 1):
        T[i, i + 1] = T[i + 1, i] = -t
    if periodic and L > 2:
        T[0, L - 1] = T[L - 1, 0] = -t
    V = np.zeros((L, L, L, L))
    for i in range(L):
        V[i, i, i, i] = U
    return T, V


def _hop_matrix(L, n, bonds):
    """One spin species: sum over bonds (i, j) of a+_i a_j + a+_j a_i on the strings of n
    electrons on L sites (Jordan-Wigner signs in site order)."""
    strs = _strings(L, n)
    idx = {s: k for k, s in enumerate(strs)}
    rows, cols, vals = [], [], []
    for k, s in enumerate(strs):
        for (i, j) in bonds:
            for (p, q) in ((i, j), (j, i)):
                if (s >> q) & 1 and not (s >> p) & 1:
                    sg = _sign(s, q)
                    s2 = s ^ (1 << q)
                    sg *= _sign(s2, p)
                    s2 |= 1 << p
                    rows.append(idx[s2])
                    cols.append(k)
                    vals.append(float(sg))
    return sp.csr_matrix((vals, (rows, cols)), shape=(len(strs), len(strs))), np.array(strs)


def hubbard_fci_energy(L, N, t=1.0, U=4.0, bonds=None, tol=1e-12):
    """Lowest eigenvalue of the Hubbard Hamiltonian -t sum_<ij>,s (a+_is a_js + h.c.) +
    U sum_i n_iu n_id in the S_z = 0 (N even) or 1/2 sector, as H_up x 1 + 1 x H_dn +
    U diag: matrix-free Lanczos over the (alpha string, beta string) grid. For a
    bipartite lattice at half filling the ground state has S = 0 (Lieb's theorem)."""
    if bonds is None:
        bonds = [(i, i + 1) for i in range(L - 1)]
    na, nb = (N + 1) // 2, N // 2
    Ha, sa = _hop_matrix(L, na, bonds)
    Hb, sb = _hop_matrix(L, nb, bonds)
    Ha, Hb = (-t * Ha).tocsr(), (-t * Hb).tocsr()
    pop = np.vectorize(lambda v: bin(int(v)).count("1"))
    dU = U * pop(np.bitwise_and.outer(sa, sb)).astype(np.float64)
    na_s, nb_s = len(sa), len(sb)

    def mv(v):
        X = v.reshape(na_s, nb_s)
        return (Ha @ X + (Hb @ X.T).T + dU * X).ravel()

    op = spla.LinearOperator((na_s * nb_s, na_s * nb_s), matvec=mv, dtype=np.float64)
    w = spla.eigsh(op, k=1, which="SA", tol=tol, v0=np.ones(na_s * nb_s), maxiter=100000)[0]
    return float(w[0])
