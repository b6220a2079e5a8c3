"""CPU tests that pin the oracle (oracle/edk_oracle.c) independently of its
own formulas: the reference's known answers, closed-form sector counts,
definition-based sectors (orbit minima + projector norms), dense linear
algebra (kron_embed, P^H O P, eigh, expm) and closed-form energies."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import oracle as orc
from tests import dense
from tests.models import Shape, blbq, hopping_pm, j1j2, rand_state, xxz

from edkit_hip.spin import spin


# ---------------------------------------------------------------- known answers
def test_gosper_known_answers():
    # abelian_overhaul_tests.jl:195-224
    s = 0b0111
    seq = []
    for _ in range(3):
        s = orc.gosper_next(s)
        seq.append(s)
    assert seq == [0b1011, 0b1101, 0b1110]
    assert orc.Basis("abelian", 4, N=2, k=0, has_k=True).D <= 6
    # basis(L, N) with no symmetry is ProjectedBasis: binomial(L, N)
    assert orc.Basis("projected", 4, N=2).D == 6
    assert orc.Basis("projected", 8, N=4).D == 70


@pytest.mark.parametrize("L", [4, 6, 8, 12, 16])
def test_benes_matches_perm(L):
    # abelian_overhaul_tests.jl:14-67: identity, reverse, cyclic and random perms
    rng = np.random.default_rng(42 + L)
    perms = [list(range(1, L + 1)), list(range(L, 0, -1)),
             [((i - 2) % L) + 1 for i in range(1, L + 1)]]
    perms += [list(rng.permutation(L) + 1) for _ in range(5)]
    states = range(2 ** L) if L <= 8 else rng.integers(0, 2 ** L, 256)
    for perm in perms:
        for s in states:
            assert orc.benes_apply(perm, L, int(s)) == orc.perm_apply_int(perm, L, int(s))


def test_apply_perm_known_answer():
    # apply_perm!([1,0,0,1], [2,3,4,1]) == [1,1,0,0] (abelian_overhaul_tests.jl:111-119)
    s = int("1001", 2)
    assert orc.perm_apply_int([2, 3, 4, 1], 4, s) == int("1100", 2)


def test_compute_period():
    L = 6
    assert orc.compute_period([((i - 2) % L) + 1 for i in range(1, L + 1)]) == L
    assert orc.compute_period(list(range(L, 0, -1))) == 2
    assert orc.compute_period(list(range(1, L + 1)), [1] * L) == 2


# ------------------------------------------------------------- sector counts
@pytest.mark.parametrize("L", [6, 8, 10])
def test_tp_small_counts(L):
    # brute-force cross-checks k0p+/k0p-/kpi p+/kpi p-/k0 at L = 6, 8, 10
    want = {6: (3, 1, 1, 3, 4), 8: (8, 2, 5, 5, 10), 10: (16, 10, 10, 16, 26)}[L]
    N = L // 2
    got = (orc.Basis("translation_parity", L, N=N, k=0, p=1).D,
           orc.Basis("translation_parity", L, N=N, k=0, p=-1).D,
           orc.Basis("translation_parity", L, N=N, k=L // 2, p=1).D,
           orc.Basis("translation_parity", L, N=N, k=L // 2, p=-1).D,
           orc.Basis("translational", L, N=N, k=0).D)
    assert got == want


@pytest.mark.parametrize("L,N", [(12, 6), (13, 5), (16, 8), (18, 7), (20, 10)])
def test_translational_dims_closed_form(L, N):
    for k in range(L):
        assert orc.Basis("translational", L, N=N, k=k).D == dense.translational_dim(L, L - N, k)


@pytest.mark.parametrize("L,N", [(12, 6), (14, 7), (16, 8), (18, 9), (20, 10), (24, 12)])
def test_tp_k0_plus_is_bracelets(L, N):
    assert orc.Basis("translation_parity", L, N=N, k=0, p=1).D == dense.bracelets_fixed(L, N)
    assert orc.Basis("translational", L, N=N, k=0).D == dense.necklaces_fixed(L, N)


def test_appendix_dims():
    # C2: L=24 N=12 k=0 necklaces; C5 spin-1/2 values
    assert orc.Basis("translational", 24, N=12, k=0).D == 112720
    assert orc.Basis("translation_parity", 24, N=12, k=0, p=1).D == 56822
    assert orc.Basis("translational", 20, N=10, k=0).D == 9252
    assert orc.Basis("translation_parity", 18, N=9, k=0, p=1).D == 1387
    # spin-1 Sz=0 k=0 (ternary necklaces with digit sum L) and its parity half
    assert orc.Basis("translational", 12, base=3, N=12, k=0).D == 6166
    assert orc.Basis("translation_parity", 12, base=3, N=12, k=0, p=1).D == 3179


# ------------------------------------------------------- tables by definition
def _check_tables(ob, L, base, gens, allowed):
    I, R = ob.tables()
    reps, norms = dense.sector_by_definition(L, base, gens, allowed)
    assert np.array_equal(I - 1, reps)
    np.testing.assert_allclose(R, norms, rtol=1e-13)
    assert np.all(np.diff(I) > 0)


@pytest.mark.parametrize("L,N", [(6, 3), (8, 4), (8, 3), (9, 4)])
def test_translational_tables_by_definition(L, N):
    allowed = lambda d: sum(d) == L - N  # noqa: E731
    for k in range(L):
        ob = orc.Basis("translational", L, N=N, k=k)
        _check_tables(ob, L, 2, [dense.translation_gens(L, k)], allowed)


@pytest.mark.parametrize("L", [6, 8, 10])
def test_tp_tables_by_definition(L):
    N = L // 2
    allowed = lambda d: sum(d) == L - N  # noqa: E731
    for k, K in ((0, 1), (L // 2, -1)):
        for p in (1, -1):
            ob = orc.Basis("translation_parity", L, N=N, k=k, p=p)
            gens = [dense.translation_gens(L, k), dense.parity_gen(L, p)]
            _check_tables(ob, L, 2, gens, allowed)


def test_abelian_tables_by_definition():
    L = 8
    for k in (0, 1, 3, 4):
        ob = orc.Basis("abelian", L, N=4, k=k, has_k=True, z=-1, has_z=True)
        gens = [dense.translation_gens(L, k), dense.flip_gen(L, -1)]
        _check_tables(ob, L, 2, gens, lambda d: sum(d) == 4)
    ob = orc.Basis("abelian", L, N=4, k=0, p=-1, has_k=True, has_p=True)
    _check_tables(ob, L, 2, [dense.translation_gens(L, 0), dense.parity_gen(L, -1)],
                  lambda d: sum(d) == 4)


def test_base3_tables_by_definition():
    L = 6
    for k in (0, 1, 3):
        ob = orc.Basis("translational", L, base=3, N=6, k=k)
        _check_tables(ob, L, 3, [dense.translation_gens(L, k)], lambda d: sum(d) == 6)
    ob = orc.Basis("translation_parity", L, base=3, N=6, k=0, p=1)
    _check_tables(ob, L, 3, [dense.translation_gens(L, 0), dense.parity_gen(L, 1)],
                  lambda d: sum(d) == 6)


@pytest.mark.parametrize("L", [6, 7, 8])
def test_parity_and_flip_tables_by_definition(L):
    """ParityBasis / FlipBasis / ParityFlipBasis / TranslationFlipBasis tables
    against orbit minima with a nonzero projected vector, and their norms
    against the projected vectors' norms (palindromes and self-flipped states
    kept only at eigenvalue +1)."""
    everything = lambda d: True  # noqa: E731
    half = lambda d: 2 * sum(d) == L  # noqa: E731
    for p in (1, -1):
        _check_tables(orc.Basis("parity", L, p=p), L, 2, [dense.parity_gen(L, p)], everything)
        _check_tables(orc.Basis("flip", L, p=p), L, 2, [dense.flip_gen(L, p)], everything)
        for z in (1, -1):
            _check_tables(orc.Basis("parity_flip", L, p=p, z=z), L, 2,
                          [dense.parity_gen(L, p), dense.flip_gen(L, z)], everything)
        if L % 2 == 0:
            N = L // 2
            _check_tables(orc.Basis("parity", L, N=N - 1, p=p), L, 2, [dense.parity_gen(L, p)],
                          lambda d: sum(d) == L - N + 1)
            _check_tables(orc.Basis("flip", L, N=N, p=p), L, 2, [dense.flip_gen(L, p)], half)
            for k in range(L):
                _check_tables(orc.Basis("translation_flip", L, N=N, k=k, p=p), L, 2,
                              [dense.translation_gens(L, k), dense.flip_gen(L, p)], half)
        for k in range(L):
            _check_tables(orc.Basis("translation_flip", L, k=k, p=p), L, 2,
                          [dense.translation_gens(L, k), dense.flip_gen(L, p)], everything)


def test_flip_bases_base3_and_unit_cell():
    L = 4
    for p in (1, -1):
        _check_tables(orc.Basis("parity", L, base=3, p=p), L, 3, [dense.parity_gen(L, p)], lambda d: True)
        _check_tables(orc.Basis("flip", L, base=3, N=L, p=p), L, 3, [dense.flip_gen(L, p)],
                      lambda d: sum(d) == L)
        for z in (1, -1):
            _check_tables(orc.Basis("parity_flip", L, base=3, p=p, z=z), L, 3,
                          [dense.parity_gen(L, p), dense.flip_gen(L, z)], lambda d: True)
        for k in range(L):
            _check_tables(orc.Basis("translation_flip", L, base=3, k=k, p=p), L, 3,
                          [dense.translation_gens(L, k), dense.flip_gen(L, p)], lambda d: True)
    L = 8
    for k in range(4):  # unit cell of two sites
        for p in (1, -1):
            _check_tables(orc.Basis("translation_flip", L, N=4, k=k, p=p, a=2), L, 2,
                          [dense.translation_gens(L, k, a=2), dense.flip_gen(L, p)],
                          lambda d: sum(d) == 4)


def test_flip_bases_reject_incompatible_arguments():
    with pytest.raises(Exception):
        orc.Basis("flip", 8, N=3, p=1)
    with pytest.raises(Exception):
        orc.Basis("parity", 8, p=2)
    with pytest.raises(Exception):
        orc.Basis("parity_flip", 8, p=1, z=0)


def test_translational_and_abelian_tables_agree():
    # basis(L, k=0).I == TranslationalBasis(L, k=0).I (abelian_overhaul_tests.jl:303-313)
    for L in (8, 10):
        a = orc.Basis("abelian", L, N=L // 2, k=0, has_k=True).tables()
        t = orc.Basis("translational", L, N=L // 2, k=0).tables()
        assert np.array_equal(a[0], t[0])
        np.testing.assert_allclose(a[1], t[1], rtol=1e-15)


# -------------------------------------------------- index() and the matvec
def _sector_matrix(op, D):
    Hs = np.zeros((D, D), np.complex128)
    for j in range(D):
        e = np.zeros(D, np.complex128)
        e[j] = 1
        Hs[:, j] = op @ e
    return Hs


CASES = [
    ("tensor", dict(), 6, 2, xxz),
    ("projected", dict(N=3), 7, 2, j1j2),
    ("translational", dict(N=4, k=0), 8, 2, xxz),
    ("translational", dict(N=4, k=1), 8, 2, j1j2),
    ("translational", dict(N=3, k=3), 8, 2, hopping_pm),
    ("translational", dict(N=4, k=2, a=2), 8, 2, xxz),
    ("translation_parity", dict(N=4, k=0, p=1), 8, 2, xxz),
    ("translation_parity", dict(N=4, k=4, p=-1), 8, 2, j1j2),
    ("abelian", dict(N=4, k=2, has_k=True), 8, 2, xxz),
    ("abelian", dict(N=4, k=0, p=-1, has_k=True, has_p=True), 8, 2, j1j2),
    ("abelian", dict(N=4, k=1, z=1, has_k=True, has_z=True), 8, 2, xxz),
    ("translational", dict(N=6, k=1), 6, 3, blbq),
    ("translation_parity", dict(N=6, k=0, p=1), 6, 3, blbq),
    ("parity", dict(N=4, p=-1), 8, 2, j1j2),
    ("parity", dict(p=1), 7, 2, xxz),
    ("flip", dict(N=4, p=1), 8, 2, xxz),
    ("flip", dict(p=-1), 7, 2, hopping_pm),
    ("parity_flip", dict(N=4, p=1, z=-1), 8, 2, j1j2),
    ("parity_flip", dict(N=4, p=-1, z=-1), 8, 2, xxz),
    ("translation_flip", dict(N=4, k=0, p=1), 8, 2, xxz),
    ("translation_flip", dict(N=4, k=3, p=-1), 8, 2, j1j2),
    ("translation_flip", dict(N=4, k=1, p=1, a=2), 8, 2, xxz),
    ("flip", dict(N=6, p=-1), 6, 3, blbq),
    ("parity_flip", dict(N=6, p=1, z=1), 6, 3, blbq),
    ("translation_flip", dict(N=6, k=2, p=-1), 6, 3, blbq),
]


@pytest.mark.parametrize("kind,kw,L,base,model", CASES)
def test_sector_matrix_is_compression_of_full_operator(kind, kw, L, base, model):
    """With P the reference's sector embedding P[s, pos] = coef(s) / order
    (TestHelpers.jl:13-23) and Q = conj(P): Q^H Q = I, H_sector = Q^H O Q and
    O Q = Q H_sector.  (index() returns the conjugate of the basis vector's
    amplitude; P itself would give the k -> -k relabelled matrix.)"""
    ob = orc.Basis(kind, L, base=base, **kw)
    H = model(Shape(L, base))
    op = orc.Op(ob, H.flatten())
    D = ob.D
    pos, coef = ob.index(np.arange(base ** L, dtype=np.uint64))
    P = np.zeros((base ** L, D), np.complex128)
    nz = coef != 0
    # order(b) is L for the dedicated types (TranslationalBasis.jl:349) while
    # the group has L/a translations: divide by the group size
    gsize = ob.order // kw.get("a", 1)
    P[np.nonzero(nz)[0], pos[nz] - 1] = coef[nz] / gsize
    Q = P.conj()
    np.testing.assert_allclose(Q.conj().T @ Q, np.eye(D), atol=1e-12)
    O = dense.full_operator(H, L, base)
    Hs = _sector_matrix(op, D)
    np.testing.assert_allclose(Q.conj().T @ O @ Q, Hs, atol=1e-12)
    np.testing.assert_allclose(O @ Q, Q @ Hs, atol=1e-12)
    np.testing.assert_allclose(Hs, Hs.conj().T, atol=1e-12)


@pytest.mark.parametrize("kind,labels", [
    ("parity", [dict(p=1), dict(p=-1)]),
    ("flip", [dict(p=1), dict(p=-1)]),
    ("parity_flip", [dict(p=p, z=z) for p in (1, -1) for z in (1, -1)]),
    ("translation_flip", [dict(k=k, p=p) for k in range(8) for p in (1, -1)]),
])
def test_spectrum_union_over_flip_sectors(kind, labels):
    """The sectors of every eigenvalue label reassemble the spectrum of the
    half-filled charge sector (nothing lost, nothing doubled)."""
    L, N = 8, 4
    H = j1j2(Shape(L), 0.35)
    O = dense.full_operator(H, L)
    I, _ = orc.Basis("projected", L, N=N).tables()
    full = np.sort(np.linalg.eigvalsh(O[np.ix_(I - 1, I - 1)]))
    vals = []
    for kw in labels:
        ob = orc.Basis(kind, L, N=N, **kw)
        if ob.D:
            vals += list(np.linalg.eigvalsh(_sector_matrix(orc.Op(ob, H.flatten()), ob.D)))
    np.testing.assert_allclose(np.sort(vals), full, atol=1e-10)


def test_spectrum_union_over_momenta():
    # basis_tests.jl:15-49 style: sector spectra reassemble the N-sector spectrum
    L, N = 8, 4
    O = dense.full_operator(j1j2(Shape(L)), L)
    pb = orc.Basis("projected", L, N=N)
    I, _ = pb.tables()
    full = np.sort(np.linalg.eigvalsh(O[np.ix_(I - 1, I - 1)]))
    vals = []
    for k in range(L):
        ob = orc.Basis("translational", L, N=N, k=k)
        vals += list(np.linalg.eigvalsh(_sector_matrix(orc.Op(ob, j1j2(Shape(L)).flatten()), ob.D)))
    np.testing.assert_allclose(np.sort(vals), full, atol=1e-12)


def test_spin_matrices_match_definitions():
    # core_tests.jl:11-21
    np.testing.assert_allclose(spin("X"), [[0, 1], [1, 0]])
    np.testing.assert_allclose(spin("Y"), [[0, -1j], [1j, 0]])
    np.testing.assert_allclose(spin("xx"), np.kron(spin("x"), spin("x")))
    sx = np.array([[0, 0.5], [0.5, 0]])
    sy = np.array([[0, -0.5j], [0.5j, 0]])
    np.testing.assert_allclose(spin("yy"), np.kron(sy, sy))
    np.testing.assert_allclose(spin((1.0, "xx"), (1.0, "yy"), (0.5, "zz")),
                               spin("xx") + spin("yy") + 0.5 * spin("zz"))
    np.testing.assert_allclose(spin("x"), sx)


def test_threaded_mul_equals_serial():
    ob = orc.Basis("translational", 16, N=8, k=3)
    op = orc.Op(ob, j1j2(Shape(16)).flatten())
    x = rand_state(ob.D, 1)
    y1 = np.zeros(ob.D, np.complex128)
    y8 = np.zeros(ob.D, np.complex128)
    op.mul_(y1, x, nthreads=1)
    op.mul_(y8, x, nthreads=8)
    np.testing.assert_allclose(y8, y1, rtol=0, atol=1e-13)
    # mul! accumulates
    op.mul_(y1, x, nthreads=1)
    np.testing.assert_allclose(y1, 2 * y8, atol=1e-12)


def test_block_mul_equals_columns():
    ob = orc.Basis("translation_parity", 12, N=6, k=0, p=1)
    op = orc.Op(ob, xxz(Shape(12), 1.5).flatten())
    X = np.stack([rand_state(ob.D, s) for s in range(3)], axis=1)
    Y = np.zeros_like(X)
    op.mul_(Y, X)
    for c in range(3):
        np.testing.assert_allclose(Y[:, c], op @ X[:, c], atol=1e-13)


def test_offdiag_count_model():
    # 2N(L-N)/(L-1) off-diagonals per row for a nearest-neighbour XX family
    # reproduces the reference's L=16 half-filling cache size 12870*(1+128/15)
    ob = orc.Basis("projected", 16, N=8)
    op = orc.Op(ob, xxz(Shape(16)).flatten())
    assert op.count_offdiag() == 12870 * 128 // 15


# ------------------------------------------------------------- eigen / Krylov
def test_tridiag_eig_matches_eigh():
    rng = np.random.default_rng(3)
    for m in (1, 2, 5, 40, 200):
        a = rng.standard_normal(m)
        b = np.abs(rng.standard_normal(max(m - 1, 0)))
        lam, Q = orc.tridiag_eig(a, b)
        T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
        np.testing.assert_allclose(lam, np.linalg.eigvalsh(T), atol=1e-12)
        np.testing.assert_allclose(T @ Q, Q * lam, atol=1e-11)
        np.testing.assert_allclose(Q.T @ Q, np.eye(m), atol=1e-12)


def test_lanczos_ground_closed_forms():
    # 4-site Heisenberg ring: E0 = -2
    ob = orc.Basis("tensor", 4)
    op = orc.Op(ob, j1j2(Shape(4), j2=0.0).flatten())
    E0, res, steps, _ = orc.lanczos_ground(op, rand_state(16, 0), 16, 1e-12)
    assert abs(E0 + 2.0) < 1e-12
    # XX ring (delta = 0) free fermions: E0 = sum of the N lowest cos(k),
    # k = 2 pi (m + phi)/L, phi = 1/2 for even N (Jordan-Wigner boundary twist)
    L, N = 12, 6
    ob = orc.Basis("projected", L, N=N)
    op = orc.Op(ob, xxz(Shape(L), 0.0).flatten())
    ks = 2 * np.pi * (np.arange(L) + 0.5) / L
    e_exact = np.sort(np.cos(ks))[:N].sum()
    E0, res, steps, _ = orc.lanczos_ground(op, rand_state(ob.D, 1), 200, 1e-10)
    assert abs(E0 - e_exact) < 1e-10


@pytest.mark.parametrize("kind,kw", [("translational", dict(N=6, k=0)),
                                     ("translation_parity", dict(N=6, k=0, p=1)),
                                     ("translational", dict(N=6, k=5))])
def test_lanczos_ground_matches_dense(kind, kw):
    ob = orc.Basis(kind, 12, **kw)
    op = orc.Op(ob, j1j2(Shape(12)).flatten())
    Hs = _sector_matrix(op, ob.D)
    e_ref = np.linalg.eigvalsh(Hs)[0]
    E0, res, steps, vec = orc.lanczos_ground(op, rand_state(ob.D, 2), 300, 1e-11, want_vec=True)
    assert abs(E0 - e_ref) < 1e-10
    assert np.linalg.norm(Hs @ vec - E0 * vec) < 1e-8


def test_krylov_dense_reference():
    # timeevolve_tests.jl:10-51: TensorBasis L=6, XXZ delta=0.7, dense exp
    L = 6
    ob = orc.Basis("tensor", L)
    H = xxz(Shape(L), 0.7)
    op = orc.Op(ob, H.flatten())
    Hd = dense.full_operator(H, L)
    psi = rand_state(ob.D, 11)
    for t in (0.1, 0.5, 1.7, 0.0):
        out, _ = orc.timeevolve(op, psi, [t], tol=1e-12, m_init=25, m_max=50)
        np.testing.assert_allclose(out[:, 0], sla.expm(-1j * t * Hd) @ psi, atol=1e-10)
    ts = np.linspace(0, 3, 31)
    out, d = orc.timeevolve(op, psi, ts, tol=1e-12, m_init=25, m_max=50)
    for i, t in enumerate(ts):
        assert np.linalg.norm(out[:, i] - sla.expm(-1j * t * Hd) @ psi) < 1e-10
    assert d["total_times_served"] == len(ts)


def test_krylov_sector():
    # timeevolve_tests.jl:53-69: TranslationalBasis L=8 N=4 k=0, +- -+ 0.5zz
    L = 8
    ob = orc.Basis("translational", L, N=4, k=0)
    op = orc.Op(ob, hopping_pm(Shape(L)).flatten())
    Hs = _sector_matrix(op, ob.D)
    psi = rand_state(ob.D, 11)
    out, _ = orc.timeevolve(op, psi, [0.3, 0.9, 1.4], tol=1e-11, m_init=20, m_max=40)
    for i, t in enumerate([0.3, 0.9, 1.4]):
        assert np.linalg.norm(out[:, i] - sla.expm(-1j * t * Hs) @ psi) < 1e-9


def test_krylov_counters():
    # timeevolve_tests.jl:71-136
    L = 6
    op = orc.Op(orc.Basis("tensor", L), xxz(Shape(L), 0.7).flatten())
    psi = rand_state(2 ** L, 5)
    _, d = orc.timeevolve(op, psi, np.linspace(0, 0.4, 25), tol=1e-10, m_init=25, m_max=50)
    assert d["basis_builds"] == 1 and d["restarts"] == 0 and d["total_times_served"] == 25
    assert d["matvecs"] <= d["max_dim_used"]
    L = 8
    H = xxz(Shape(L), 0.7)
    op = orc.Op(orc.Basis("tensor", L), H.flatten())
    Hd = dense.full_operator(H, L)
    psi = rand_state(2 ** L, 6)
    ts = np.linspace(0, 20, 11)
    out, d = orc.timeevolve(op, psi, ts, tol=1e-10, m_init=6, m_max=6, extend_basis=False)
    for i, t in enumerate(ts):
        assert np.linalg.norm(out[:, i] - sla.expm(-1j * t * Hd) @ psi) < 1e-6
    assert d["restarts"] >= 1 and d["basis_builds"] >= 2
    L = 6
    op = orc.Op(orc.Basis("tensor", L), xxz(Shape(L), 0.7).flatten())
    _, d = orc.timeevolve(op, rand_state(2 ** L, 7), np.linspace(0, 3, 15), tol=1e-12,
                          m_init=5, m_max=40, extend_step=5)
    assert d["basis_extensions"] >= 1


def test_krylov_errors():
    # timeevolve_tests.jl:138-201 error paths
    op = orc.Op(orc.Basis("tensor", 4), xxz(Shape(4)).flatten())
    with pytest.raises(ValueError):
        orc.Krylov(op, np.zeros(16, np.complex128))
    c = orc.Krylov(op, rand_state(16, 1), m_init=4, m_max=8)
    c.evolve(1.0)
    with pytest.raises(ValueError):
        c.evolve(0.5)
    with pytest.raises(ValueError):
        orc.timeevolve(op, rand_state(16, 1), [-1.0])


# ------------------------------------------------------------- schmidt / ent_S
SCHMIDT_CASES = [
    ("tensor", dict(), 7, 2),
    ("projected", dict(N=4), 8, 2),
    ("translational", dict(N=4, k=3), 8, 2),
    ("translational", dict(N=4, k=1, a=2), 8, 2),
    ("translation_parity", dict(N=4, k=4, p=-1), 8, 2),
    ("parity", dict(N=4, p=-1), 8, 2),
    ("flip", dict(N=4, p=1), 8, 2),
    ("parity_flip", dict(N=4, p=-1, z=1), 8, 2),
    ("translation_flip", dict(N=4, k=3, p=-1), 8, 2),
    ("abelian", dict(N=4, k=2, has_k=True, z=-1, has_z=True), 8, 2),
    ("translational", dict(N=6, k=1), 6, 3),
]


@pytest.mark.parametrize("kind,kw,L,base", SCHMIDT_CASES)
def test_schmidt_matrix_is_the_embedded_state(kind, kw, L, base):
    """schmidt(v, Ainds, b) element by element against the full-space state
    Q v built from index() (the embedding of the compression test above),
    reshaped by the A and B digits; a unit vector gives a unit matrix."""
    ob = orc.Basis(kind, L, base=base, **kw)
    rng = np.random.default_rng(3)
    v = rng.normal(size=ob.D) + 1j * rng.normal(size=ob.D)
    v /= np.linalg.norm(v)
    A = [1, 2, 4, L - 1]
    M = ob.schmidt(v, A)
    n = base ** L
    pos, coef = ob.index(np.arange(n, dtype=np.uint64))
    gsize = ob.order // kw.get("a", 1) if kind != "abelian" else ob.order
    gsize = 1 if kind in ("tensor", "projected") else gsize
    P = np.zeros((n, ob.D), np.complex128)
    nz = coef != 0
    P[np.nonzero(nz)[0], pos[nz] - 1] = coef[nz] / gsize
    psi = P.conj() @ v
    Bs = [q for q in range(1, L + 1) if q not in A]
    Md = np.zeros_like(M)
    for c in range(n):
        d = dense.digits(c, L, base)
        Md[dense.content([d[q - 1] for q in A], base), dense.content([d[q - 1] for q in Bs], base)] = psi[c]
    np.testing.assert_allclose(M, Md, atol=1e-14)
    assert abs(np.linalg.norm(M) - 1) < 1e-12
