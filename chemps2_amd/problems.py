"""Hamiltonian generators and the FCIDUMP reader.

Integrals are returned in the physicist convention V[i, j, k, l] = <ij|kl> = (ik|jl)
with orbital irreps in CheMPS2's XOR labelling (Irreps.h), the layout that
``Hamiltonian::setVmat`` receives (Hamiltonian.cpp:296-423 for the FCIDUMP reader).
"""
import numpy as np

def _molpro_from_xor(group):
    """XOR label -> Molpro irrep number, per point group (Irreps.cpp psi2molpro tables)."""
    return {
        0: [1],
        1: [1, 2], 2: [1, 2], 3: [1, 2],
        4: [1, 4, 3, 2],
        5: [1, 4, 2, 3], 6: [1, 4, 2, 3],
        7: [1, 4, 6, 7, 8, 5, 3, 2],
    }[group]


def molpro_to_xor(group, molpro_irrep):
    """Map a Molpro irrep number (1-based) to CheMPS2's XOR label."""
    return _molpro_from_xor(group).index(molpro_irrep)


def hubbard(L, t=1.0, U=4.0, periodic=False):
    """1D Hubbard chain: T[i,i+1] = -t, <ii|ii> = U."""
    T = np.zeros((L, L))
    for i in range(L - 1):
        T[i, i + 1] = T[i + 1, i] = -t
    if periodic and L > 2:
        T[0, L - 1] = T[L - 1, 0] = -t
    V = np.zeros((L, L, L, L))
    for i in range(L):
        V[i, i, i, i] = U
    return T, V


def hubbard_cylinder(nx=8, ny=4, t=1.0, U=8.0):
    """2D Hubbard model on an nx x ny cylinder, periodic around the ny-site circumference
    and open along nx, snaked into a chain: site (x, y) has index ny x + y for even x and
    ny x + ny - 1 - y for odd x. Returns T (hopping -t on every bond), V (U on site) and
    the list of bonds (i, j) in chain indices."""
    L = nx * ny

    def idx(x, y):
        return ny * x + (y if x % 2 == 0 else ny - 1 - y)

    bonds = set()
    for x in range(nx):
        for y in range(ny):
            if ny > 2 or y + 1 < ny:
                bonds.add(tuple(sorted((idx(x, y), idx(x, (y + 1) % ny)))))
            if x + 1 < nx:
                bonds.add(tuple(sorted((idx(x, y), idx(x + 1, y)))))
    T = np.zeros((L, L))
    for i, j in bonds:
        T[i, j] = T[j, i] = -t
    V = np.zeros((L, L, L, L))
    for i in range(L):
        V[i, i, i, i] = U
    return T, V, sorted(bonds)


def random_abinitio(L, seed, group=7):
    """Seeded synthetic ab-initio Hamiltonian with D2h (or smaller) symmetry.

    Orbital irreps uniform in {0..nirrep-1}; h_ij ~ N(0, (0.1 exp(-|i-j|/4))^2) for
    equal irreps, h_ii ~ U[-2, 0]; (ij|kl) = sum_P B^P_ij B^P_kl over 3L symmetric
    factors with B^P_ij ~ N(0, (0.05 exp(-|i-j|/6))^2) masked to irrep_i ^ irrep_j =
    Gamma_P. This gives a positive semidefinite ERI tensor with 8-fold symmetry and
    the selection rule irrep_i ^ irrep_j = irrep_k ^ irrep_l.
    Returns (irreps, T, V).
    """
    nirrep = {0: 1, 1: 2, 2: 2, 3: 2, 4: 4, 5: 4, 6: 4, 7: 8}[group]
    rng = np.random.default_rng(seed)
    irreps = rng.integers(0, nirrep, size=L).astype(np.int32)
    idx = np.arange(L)
    dist = np.abs(idx[:, None] - idx[None, :])
    same = irreps[:, None] == irreps[None, :]
    h = rng.normal(size=(L, L)) * 0.1 * np.exp(-dist / 4.0)
    h = np.triu(h, 1)
    h = h + h.T
    h[~same] = 0.0
    h[idx, idx] = rng.uniform(-2.0, 0.0, size=L)
    eri = np.zeros((L, L, L, L))
    xor = irreps[:, None] ^ irreps[None, :]
    for _ in range(3 * L):
        g = rng.integers(0, nirrep)
        B = rng.normal(size=(L, L)) * 0.05 * np.exp(-dist / 6.0)
        B = np.triu(B) + np.triu(B, 1).T
        B[xor != g] = 0.0
        eri += np.einsum("ij,kl->ijkl", B, B)
    V = np.ascontiguousarray(eri.transpose(0, 2, 1, 3))  # (ij|kl) -> <ik|jl>
    return irreps, h, V


def mx_table(T, V, N):
    """The L^4 table h_{ab;cd} = V_abcd + (T_ac delta_bd + delta_ac T_bd) / (N - 1) that
    Problem::construct_mxelem builds (Problem.cpp:363-384), flattened as
    mx[a + L(b + L(c + L d))]."""
    T = np.asarray(T, dtype=np.float64)
    L = T.shape[0]
    V = np.asarray(V, dtype=np.float64).reshape(L, L, L, L)
    eye = np.eye(L)
    mx = V + (np.einsum("ac,bd->abcd", T, eye) + np.einsum("ac,bd->abcd", eye, T)) / (N - 1)
    return np.ascontiguousarray(mx.transpose(3, 2, 1, 0)).ravel()


def reorder_d2h(irreps, group=7):
    """DMRG orbital order for D2h: irreps grouped as Ag B1u B3u B2g B2u B3g B1g Au
    (Problem::SetupReorderD2h, Problem.cpp:57-90). Returns f2 (DMRG -> Hamiltonian index);
    the identity for other groups."""
    L = len(irreps)
    if group != 7:
        return list(range(L))
    order = [0, 5, 7, 2, 6, 3, 1, 4]
    return [h for o in order for h in range(L) if irreps[h] == o]


def permute(irreps, T, V, f2):
    """Irreps, T and V in the orbital order f2 (new index -> old index)."""
    f2 = np.asarray(f2)
    irr = np.asarray(irreps)[f2]
    return irr, np.asarray(T)[np.ix_(f2, f2)], np.asarray(V)[np.ix_(f2, f2, f2, f2)]


def read_fcidump(path):
    """Reads an FCIDUMP file (Hamiltonian.cpp:296-423). Returns a dict with L, N, TwoS,
    irreps (XOR labels, D2h subgroup given by ``group``), T, V (physicist), econst."""
    with open(path) as f:
        text = f.read()
    head, _, body = text.partition("&END")
    if not body:
        head, _, body = text.partition("/")
    hd = head.replace("\n", " ").replace(",", " ")
    import re
    L = int(re.search(r"NORB\s*=\s*(\d+)", hd).group(1))
    N = int(re.search(r"NELEC\s*=\s*(\d+)", hd).group(1))
    ms2 = re.search(r"MS2\s*=\s*(\d+)", hd)
    TwoS = int(ms2.group(1)) if ms2 else 0
    orbsym = re.search(r"ORBSYM\s*=\s*([\d\s]+)", hd)
    osym = [int(x) for x in orbsym.group(1).split()][:L] if orbsym else [1] * L
    T = np.zeros((L, L))
    V = np.zeros((L, L, L, L))
    econst = 0.0
    for line in body.strip().splitlines():
        parts = line.split()
        if len(parts) != 5:
            continue
        val = float(parts[0].replace("D", "E").replace("d", "e"))
        i, j, k, l = (int(x) for x in parts[1:])
        if i == j == k == l == 0:
            econst = val
        elif k == 0 and l == 0:
            T[i - 1, j - 1] = T[j - 1, i - 1] = val
        else:
            # chemist (ij|kl) -> physicist <ik|jl>, all 8 permutations
            a, b, c, d = i - 1, k - 1, j - 1, l - 1  # <ab|cd> = (ac|bd)
            for (p, q, r, s) in ((a, b, c, d), (b, a, d, c), (c, b, a, d), (a, d, c, b),
                                 (c, d, a, b), (d, c, b, a), (b, c, d, a), (d, a, b, c)):
                V[p, q, r, s] = val
    return {"L": L, "N": N, "TwoS": TwoS, "molpro_irreps": osym, "T": T, "V": V, "econst": econst}
