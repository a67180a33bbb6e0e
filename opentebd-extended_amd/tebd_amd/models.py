"""Bond Hamiltonians, Trotter gates and product states in the reference's conventions.

Host-side set-up (d^2 x d^2 matrices, negligible cost), kept on the host as in
the reference.  Conventions:

* Kronecker products: tensorProd(A,B)((i-1)dB+j,(k-1)dB+l) = A(i,k) B(j,l),
  left site slow (VidalCore_v_2.0/TEBDtools_module.f90:253-322) == numpy.kron.
* One-site terms are split half/half over the two bonds that touch a site, with
  full weight on the edge bonds (HamiOneSite/HamiLeft/HamiRight,
  HamiltonianCore_v_2.0/Hamiltonian_tools_module.f90:816-886).
* Gates: U = V exp(-i t E) V^dagger from the Hermitian eigendecomposition of
  H (matrix_exponential_c, TEBDtools_module.f90:176-222); 2nd-order OBC
  schedule: odd bonds t = dt/2, even bonds t = dt (NEWConstructPropagators,
  TEBDtools_module.f90:1079-1101).  Imaginary time passes t = -i dtau
  (propagation_module.f90:656-662).
"""
import numpy as np

# ---------------------------------------------------------------------------
# local operators
# ---------------------------------------------------------------------------


def spin_ops(spin):
    """Sx, Sy, Sz of spin S in the basis Sz = S - k, k = 0..2S
    (CreateHeisenbergOps, heisenberg_module.f90:52-118)."""
    d = int(round(2 * spin + 1))
    m = spin - np.arange(d)
    c = np.sqrt((spin + m[:-1]) * (spin + 1 - m[:-1]))  # <m|S+|m-1> on the superdiagonal
    sp = np.zeros((d, d))
    sp[np.arange(d - 1), np.arange(1, d)] = c
    sx = 0.5 * (sp + sp.T)
    sy = 0.5j * (sp.T - sp)
    sz = np.diag(m)
    return sx, sy.astype(np.complex128), sz


def boson_ops(max_filling):
    """Destruction operator a (a(j,j+1) = sqrt(j)) and number operator, d = maxFilling+1
    (CreateFieldOps, bose_hubbard_module.f90:50-111)."""
    d = max_filling + 1
    a = np.zeros((d, d))
    for j in range(1, d):
        a[j - 1, j] = np.sqrt(j)
    return a, a.T @ a


def fermion_ops(max_filling=2):
    """Spin-1/2 fermions with at most max_filling particles per site.

    Basis from onsiteStateListIdof with fermi statistics
    (Hamiltonian_tools_module.f90:205-367): ordered by particle number, then by
    the component holding the particle; for maxFilling=2 that is
    [vacuum, up, down, up+down].  Destruction operators carry the intra-site
    Jordan-Wigner phase (-1)^{sum_{l<k} n_l} (fermi_hubbard_module.f90:185-199).
    Returns (a_up, a_dn, n_tot, Sz, F) with F = diag((-1)^n) the fermi phase operator.
    """
    states = [(0, 0)]
    if max_filling >= 1:
        states += [(1, 0), (0, 1)]
    if max_filling >= 2:
        states += [(1, 1)]
    d = len(states)
    idx = {s: i for i, s in enumerate(states)}
    a = [np.zeros((d, d)), np.zeros((d, d))]
    for k in range(2):
        for j, sj in enumerate(states):
            if sj[k] == 0:
                continue
            new = list(sj)
            new[k] -= 1
            phase = (-1.0) ** sum(sj[:k])
            a[k][idx[tuple(new)], j] = phase
    ntot = a[0].T @ a[0] + a[1].T @ a[1]
    sz = 0.5 * (a[0].T @ a[0]) - 0.5 * (a[1].T @ a[1])
    F = np.diag((-1.0) ** np.round(np.diag(ntot)))
    return a[0], a[1], ntot, sz, F


# ---------------------------------------------------------------------------
# bond Hamiltonians
# ---------------------------------------------------------------------------


def heisenberg(L, spin=0.5, Jx=1.0, Jy=1.0, Jz=1.0, field=None, bc="O"):
    """Jx SxSx + Jy SySy + Jz SzSz per bond (HamiltonianHeisenberg,
    heisenberg_module.f90:143-174) plus a site-dependent field h_i Sz_i split
    over bonds the way the reference splits its uniform field (:158-167).  With
    periodic boundaries (bc='P') a last entry H(L) couples site L to site 1, with the
    coupling only (:169-171)."""
    sx, sy, sz = spin_ops(spin)
    d = sx.shape[0]
    one = np.eye(d)
    base = Jx * np.kron(sx, sx) + Jy * np.kron(sy, sy) + Jz * np.kron(sz, sz)
    H = np.array([base.astype(np.complex128) for _ in range(L - 1)])
    if field is not None:
        h = np.asarray(field, float)
        left, right = 0.5 * np.kron(sz, one), 0.5 * np.kron(one, sz)
        for b in range(L - 1):
            H[b] += h[b] * left + h[b + 1] * right
        H[0] += h[0] * left
        H[L - 2] += h[L - 1] * right
    if bc == "P":
        H = np.concatenate([H, base.astype(np.complex128)[None]], axis=0)
    return H


def bose_hubbard(L, max_filling, J, U0, mu0=0.0, V0=0.0):
    """Extended Bose-Hubbard model with bond-dependent hopping J_b
    (HamiltonianBoseHubbardJVector, bose_hubbard_module.f90:275-345)."""
    a, n = boson_ops(max_filling)
    d = a.shape[0]
    one = np.eye(d)
    J = np.broadcast_to(np.asarray(J, float), (L - 1,))
    t_op = np.kron(a.T, a) + np.kron(a.T, a).T
    nn = n @ n
    onsite_l = 0.25 * U0 * (np.kron(nn, one) - np.kron(n, one)) - 0.5 * mu0 * np.kron(n, one)
    onsite_r = 0.25 * U0 * (np.kron(one, nn) - np.kron(one, n)) - 0.5 * mu0 * np.kron(one, n)
    H = np.zeros((L - 1, d * d, d * d), np.complex128)
    for b in range(L - 1):
        H[b] = onsite_l + onsite_r - J[b] * t_op + V0 * np.kron(n, n)
    H[0] += onsite_l
    H[L - 2] += onsite_r
    return H


def fermi_hubbard(L, t=1.0, U0=4.0, mu0=0.0, max_filling=2):
    """Spin-1/2 Hubbard chain, U (n_up - 1/2)(n_dn - 1/2) on-site form
    (HamiltonianHubbard, fermi_hubbard_module.f90:289-356; tunnelling ttot with
    the fermi phase, :236-241)."""
    aup, adn, ntot, _, F = fermion_ops(max_filling)
    d = aup.shape[0]
    one = np.eye(d)
    ttot = np.zeros((d * d, d * d))
    for a in (aup, adn):
        x = np.kron(a, one)
        y = np.kron(F, a)
        ttot += x.T @ y + y.T @ x
    hub = (aup.T @ aup - 0.5 * one) @ (adn.T @ adn - 0.5 * one)
    H = np.zeros((L - 1, d * d, d * d), np.complex128)
    for b in range(L - 1):
        H[b] = (-0.5 * mu0) * (np.kron(ntot, one) + np.kron(one, ntot)) - t * ttot \
            + 0.5 * U0 * (np.kron(hub, one) + np.kron(one, hub))
    H[0] += (-0.5 * mu0) * np.kron(ntot, one) + 0.5 * U0 * np.kron(hub, one)
    H[L - 2] += (-0.5 * mu0) * np.kron(one, ntot) + 0.5 * U0 * np.kron(one, hub)
    return H


# ---------------------------------------------------------------------------
# gates
# ---------------------------------------------------------------------------


def matrix_exponential(H, t):
    """exp(-i t H) of a Hermitian matrix through its eigendecomposition
    (matrix_exponential_c, TEBDtools_module.f90:176-222); t may be complex."""
    E, V = np.linalg.eigh(H, UPLO="L")
    return (V * np.exp(-1j * t * E)[None, :]) @ V.conj().T


# Forest-Ruth theta as the reference evaluates it, 1 / (2 - 2**(1.0/3.0)) with a
# single-precision exponent 1.0/3.0 (TEBDtools_module.f90:1135): 1.3512072077991604
# instead of 1.3512071919596578.
FR_THETA = 1.0 / (2.0 - 2.0 ** float(np.float32(1.0) / np.float32(3.0)))


def gate_table(L, order=2, bc="O"):
    """The propagator table of NEWConstructPropagators (TEBDtools_module.f90:1104-1193)
    as a list of (bond, time factor) per entry, bond 1-based (L = the edge bond between
    sites L and 1), so that entry k is exp(-i factor dt H(bond)).  Its length is
    AllocateProp's (:937-998): L-1 (2nd order, open), L (2nd, periodic), 2(L-1) (5th,
    open), 2L or 3L (5th, periodic, L even or odd)."""
    th = FR_THETA
    n = L - 1
    if order == 2 and bc == "O":
        return [(i, 0.5 if i % 2 == 1 else 1.0) for i in range(1, L)]
    if order == 2 and bc == "P":
        even = 1.0 if L % 2 == 0 else 0.5
        return [(i, 0.5 if i % 2 == 1 else even) for i in range(1, L)] + [(L, 1.0)]
    if order != 5:
        raise ValueError("trotter order must be 2 or 5")
    if bc == "O" or L % 2 == 0:
        size = 2 * n if bc == "O" else 2 * L
        tab = [None] * size
        for i in range(1, L):
            f = (0.5 * th, 0.5 * (1.0 - th)) if i % 2 == 1 else (th, 1.0 - 2.0 * th)
            tab[2 * (i - 1)], tab[2 * (i - 1) + 1] = (i, f[0]), (i, f[1])
        if bc == "P":
            tab[2 * n], tab[2 * n + 1] = (L, th), (L, 1.0 - 2.0 * th)
        return tab
    tab = [None] * (3 * L)
    for i in range(1, L, 2):
        tab[3 * (i - 1)], tab[3 * (i - 1) + 1] = (i, 0.5 * th), (i, 0.5 * (1.0 - th))
    for i in range(2, L, 2):
        b = 2 + 3 * (i - 2)
        tab[b:b + 4] = [(i, 0.5 * th * th), (i, 0.5 * (1.0 - th) * th), (i, 0.5 * (1.0 - 2.0 * th) * th),
                        (i, 0.5 * (1.0 - th) * (1.0 - 2.0 * th))]
    tab[3 * n:3 * n + 3] = [(L, th * th), (L, (1.0 - 2.0 * th) * th), (L, (1.0 - 2.0 * th) ** 2)]
    return tab


def construct_propagators(H, dt, order=2, bc="O"):
    """Trotter gates of the reference's schedules (NEWConstructPropagators,
    TEBDtools_module.f90:1079-1193): entry k of gate_table is exp(-i factor dt H(bond)).
    2nd order, open boundaries: odd bonds (1-based) exp(-i dt/2 H), even bonds
    exp(-i dt H).  H: L-1 bond Hamiltonians (open) or L (periodic, H[L-1] the edge)."""
    H = np.asarray(H)
    L = H.shape[0] + 1 if bc == "O" else H.shape[0]
    tab = gate_table(L, order, bc)
    U = np.empty((len(tab),) + H.shape[1:], dtype=np.complex128)
    for k, (b, f) in enumerate(tab):
        U[k] = matrix_exponential(H[b - 1], f * dt)
    return U


def read_propagator(real_path, imag_path):
    """A gate from two list-directed text files of its real and imaginary parts, read as
    Fortran reads a whole n x n array (column-major), the way the Extensible wrapper loads
    its Propagator_*.nml files (Case_Studies/Extensible_Wrapper/TEBDSolve.f90:172-206)."""
    re = np.array(open(real_path).read().replace(",", " ").split(), float)
    im = np.array(open(imag_path).read().replace(",", " ").split(), float)
    n = int(round(np.sqrt(re.size)))
    if n * n != re.size or im.size != re.size:
        raise ValueError(f"{real_path}, {imag_path}: not the two parts of one square matrix")
    return re.reshape(n, n, order="F") + 1j * im.reshape(n, n, order="F")


def file_gate_table(L, left, bulk_odd, bulk_even, right):
    """The gate table TEBDSolve.f90:172-206 assembles from files: bond 1 `left`, odd bonds
    3, 5, .. <= L-2 `bulk_odd`, even bonds 2, 4, .. <= L-2 `bulk_even`, bond L-1 `right`
    (assigned last, so for odd L it replaces the even-bond gate there)."""
    U = np.empty((L - 1,) + left.shape, np.complex128)
    U[0] = left
    for i in range(3, L - 1, 2):
        U[i - 1] = bulk_odd
    for i in range(2, L - 1, 2):
        U[i - 1] = bulk_even
    U[L - 2] = right
    return U


# ---------------------------------------------------------------------------
# initial states (reference padded layout: gam [site, b, i, a], lam [bond, a])
# ---------------------------------------------------------------------------


def product_state(coeffs, chi):
    """ProductStateMPD (Hamiltonian_tools_module.f90:89-121): Gamma(1,j,1) = c_j, lambda = (1,0,..)."""
    c = np.asarray(coeffs, np.complex128)
    L, d = c.shape
    gam = np.zeros((L, chi, d, chi), np.complex128)
    gam[:, 0, :, 0] = c
    lam = np.zeros((L + 1, chi))
    lam[:, 0] = 1.0
    return gam, lam


def all_states(L, d, chi):
    """AllStates (Hamiltonian_tools_module.f90:181-202): every local state with weight 1/sqrt(d)."""
    return product_state(np.full((L, d), 1.0 / np.sqrt(d)), chi)


def neel(L, d, up, down):
    c = np.zeros((L, d), np.complex128)
    for s in range(L):
        c[s, up if s % 2 == 0 else down] = 1.0
    return c


def product_labels(coeffs, chi, conserv=None, tot_num=None):
    """Labels of a product state of number eigenstates for the number-conserving update
    (ProductStateLabels, Hamiltonian_tools_module.f90:124-178): the number left of each
    link in its one Schmidt state, from each site's nonzero coefficient, the right labels
    tot_num - left, 10000 for every other state.  conserv: the number of each local state
    (default: its index); tot_num defaults to the state's total."""
    c = np.asarray(coeffs)
    L, d = c.shape
    cv = np.arange(d) if conserv is None else np.asarray(conserv)
    LL = np.full((L + 1, chi), 10000, np.int32)
    LR = np.full((L + 1, chi), 10000, np.int32)
    LL[0, 0] = 0
    for s in range(L):
        nz = np.flatnonzero(np.abs(c[s]) != 0.0)
        if nz.size == 0 or np.any(cv[nz] != cv[nz[0]]):
            raise ValueError(f"site {s + 1}: not a number eigenstate")
        LL[s + 1, 0] = LL[s, 0] + int(cv[nz[0]])
    tot = int(LL[L, 0]) if tot_num is None else int(tot_num)
    LR[:, 0] = tot - LL[:, 0]
    return LL, LR

