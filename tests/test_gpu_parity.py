"""GPU parity: the engine's kernels, operators and whole solves through the
C-ABI (libbbjd.so) against the oracle (CPU restatement of the reference,
oracle/) on the same seeded inputs.

Floating-point bar: orthogonalisation / operator outputs agree to a few ulp
relative to the operand norms (1e-12); converged eigenvalues agree with the
oracle's to 1e-8 relative (the parity bar of the reference tests, e.g.
tests/test_bbjd.py:276) and every residual is below tol.
"""
import ctypes as C

import numpy as np
import pytest

from bbjd_amd import BandedTauPencil, DeviceOperator, _lib, jdqz
from bbjd_amd._bbjd_c import jdqz_c
from oracle import orc
from tests.helpers import make_diagonal, match_error, rel_match, residual

pytestmark = pytest.mark.gpu


def _orthonormal(n, k, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    q, _ = np.linalg.qr(M)
    return np.asfortranarray(q)


@pytest.mark.parametrize("n,k", [(3000, 17), (2049, 1), (70000, 40), (5, 3)])
def test_cgs2_matches_oracle(n, k):
    lib = _lib.load()
    olib = orc.load_oracle()
    rng = np.random.default_rng(n + k)
    Q = _orthonormal(n, k, n)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for case in ("generic", "near_parallel"):
        x = v.copy() if case == "generic" else Q @ rng.standard_normal(k) + 1e-6 * v
        vg, vo = x.copy(), x.copy()
        hg = np.zeros(k, np.complex128)
        ho = np.zeros(k, np.complex128)
        ng = lib.bbjd_block_orthogonalize(n, k, Q.ctypes.data, vg.ctypes.data, hg.ctypes.data)
        no = olib.orc_block_orthogonalize(n, k, Q.ctypes.data, vo.ctypes.data, ho.ctypes.data)
        scale = np.linalg.norm(x)
        assert np.linalg.norm(vg - vo) <= 1e-12 * scale, case
        assert abs(ng - no) <= 1e-12 * scale, case
        assert abs(ng - np.linalg.norm(vg)) <= 1e-12 * scale
        # the output is orthogonal to Q to rounding
        assert np.linalg.norm(Q.conj().T @ vg) <= 1e-12 * scale


@pytest.mark.parametrize("proj_blocks,rows", [(8, 64), (256, 128), (1024, 256), (6000, 512), (100, 1024)])
def test_cgs2_every_row_tile(proj_blocks, rows):
    """each row tile / chunking of the streaming kernels (forced), fused and
    two-launch finishing, against the oracle; ragged last tiles"""
    lib = _lib.load()
    olib = orc.load_oracle()
    try:
        for n, k, fuse in ((40000 + 77, 33, 48), (3000 + 5, 9, 48), (3000 + 5, 9, 0), (3000 + 5, 9, 16)):
            assert lib.bbjd_set_tiles(proj_blocks, rows, 0) == 0
            lib.bbjd_set_fuse_blocks(fuse)
            rng = np.random.default_rng(n + k + rows)
            Q = _orthonormal(n, k, n + 1)
            x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
            vg, vo = x.copy(), x.copy()
            hg = np.zeros(k, np.complex128)
            ho = np.zeros(k, np.complex128)
            ng = lib.bbjd_block_orthogonalize(n, k, Q.ctypes.data, vg.ctypes.data, hg.ctypes.data)
            no = olib.orc_block_orthogonalize(n, k, Q.ctypes.data, vo.ctypes.data, ho.ctypes.data)
            scale = np.linalg.norm(x)
            assert np.linalg.norm(vg - vo) <= 1e-12 * scale, (n, k, fuse)
            assert np.linalg.norm(hg - ho) <= 1e-12 * scale, (n, k, fuse)
            assert abs(ng - no) <= 1e-12 * scale, (n, k, fuse)
    finally:
        lib.bbjd_set_tiles(192, 0, 2048)
        lib.bbjd_set_fuse_blocks(16)


FUSED_K = [1, 2, 7, 8, 9, 16, 17, 31, 32, 33, 64, 65, 100, 128, 129]


@pytest.mark.parametrize("n", [1000 + 3, 70000 + 5])
def test_fused_cgs2_matches_oracle(n):
    """the three-read CGS2 (first update and second projection in one pass
    over Q, k <= 128) and the four-launch form against the oracle: every
    columns-per-wave width, the k = 129 fallback, both DGKS branches, the
    Hessenberg sum h1 + h2, unit-norm output (mode 2 forces the route where
    the second pass leaves its result to a scaling by the computed norm) and
    the two-pass projection"""
    lib = _lib.load()
    olib = orc.load_oracle()
    rng = np.random.default_rng(n)
    Qall = _orthonormal(n, max(FUSED_K), n + 7)
    flags = np.zeros(8, np.int32)
    try:
        for k in FUSED_K:
            Q = Qall[:, :k]
            assert Q.flags.f_contiguous
            v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
            r = rng.standard_normal(k) + 1j * rng.standard_normal(k)
            cases = {"generic": v, "near_parallel": Q @ r + 1e-6 * v, "nearer": Q @ r + 1e-11 * v}
            for case, x in cases.items():
                scale = np.linalg.norm(x)
                vo, ho = x.copy(), np.zeros(k, np.complex128)
                no = olib.orc_block_orthogonalize(n, k, Q.ctypes.data, vo.ctypes.data, ho.ctypes.data)
                tol_unit = 1e-12 * scale / no
                for mode in (1, 2, 0):
                    lib.bbjd_set_cgs_fused(mode)
                    tag = (k, case, mode)
                    vg, hg = x.copy(), np.zeros(k, np.complex128)
                    ng = lib.bbjd_block_orthogonalize(n, k, Q.ctypes.data, vg.ctypes.data, hg.ctypes.data)
                    assert np.linalg.norm(vg - vo) <= 1e-12 * scale, tag
                    # h_work: the reference leaves the last pass's coefficients
                    # there, the engine the sum of both passes (the Hessenberg
                    # column of its GMRES), which is Q^H x to rounding
                    href = ho if case == "generic" else Q.conj().T @ x
                    assert np.linalg.norm(hg - href) <= 1e-12 * scale, tag
                    assert abs(ng - no) <= 1e-12 * scale, tag
                    vu = x.copy()
                    nu = lib.bbjd_block_orthonormalize(n, k, Q.ctypes.data, vu.ctypes.data, flags.ctypes.data)
                    assert flags[0] == (case != "generic"), tag      # the DGKS branch taken
                    assert nu == ng, tag            # the same reductions: the norm before scaling
                    assert abs(np.linalg.norm(vu) - 1.0) <= 1e-13, tag
                    assert np.linalg.norm(vu - vo / no) <= tol_unit, tag
                    assert np.linalg.norm(Q.conj().T @ vu) <= tol_unit, tag
                    vp, vq = x.copy(), x.copy()
                    assert lib.bbjd_project_out_host(n, k, Q.ctypes.data, vp.ctypes.data) == 0
                    olib.orc_project_out(n, k, Q.ctypes.data, vq.ctypes.data)
                    assert np.linalg.norm(vp - vq) <= 1e-12 * scale, tag
    finally:
        lib.bbjd_set_cgs_fused(1)


@pytest.mark.parametrize("n,k", [(1003, 1), (1003, 9), (70005, 33), (70005, 100), (20001, 128), (5000, 130)])
def test_test_space_expansion_pencil_column(n, k):
    """the expansion's test-space step: w = nu0 Av + mu0 Bv formed in the
    first passes over W and the pencil column [W w]^H [Av Bv] taken in the
    first pass (k <= 128), or the combination, CGS2 and projections one after
    another (k = 130); the column against numpy, w against the oracle's CGS2
    of the same combination, normalised; both DGKS branches"""
    lib = _lib.load()
    olib = orc.load_oracle()
    rng = np.random.default_rng(3 * n + k)
    W = _orthonormal(n, k, n + k + 5)
    sigma = 0.7 - 1.3j
    nu0 = 1 / np.sqrt(1 + abs(sigma) ** 2)
    mu0 = -sigma * nu0
    for case in ("generic", "near_parallel"):
        a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        if case == "near_parallel":
            a = W @ (rng.standard_normal(k) + 1j * rng.standard_normal(k)) + 1e-7 * a
            b = W @ (rng.standard_normal(k) + 1j * rng.standard_normal(k)) + 1e-7 * b
        w0 = nu0 * a + mu0 * b
        wo, ho = w0.copy(), np.zeros(k, np.complex128)
        no = olib.orc_block_orthogonalize(n, k, W.ctypes.data, wo.ctypes.data, ho.ctypes.data)
        wg = np.zeros(n, np.complex128)
        cA = np.zeros(k + 1, np.complex128)
        cB = np.zeros(k + 1, np.complex128)
        ng = lib.bbjd_expand_test_space(n, k, W.ctypes.data, _lib.zc(nu0), _lib.zc(mu0), a.ctypes.data,
                                        b.ctypes.data, wg.ctypes.data, cA.ctypes.data, cB.ctypes.data)
        scale = np.linalg.norm(w0)
        assert abs(ng - no) <= 1e-12 * scale, case
        assert abs(np.linalg.norm(wg) - 1.0) <= 1e-13, case
        assert np.linalg.norm(wg - wo / no) <= 1e-12 * scale / no, case
        Wx = np.column_stack([W, wg])
        assert np.linalg.norm(cA - Wx.conj().T @ a) <= 1e-12 * np.linalg.norm(a), case
        assert np.linalg.norm(cB - Wx.conj().T @ b) <= 1e-12 * np.linalg.norm(b), case


# super-block sizes hw * bs = 4, 16, 8, 8 (compiled loops) and 12 (generic)
SHAPES = [(2000, 4, 1, 2), (4096, 8, 2, 16), (1000, 8, 1, 4), (640, 4, 2, 0), (1206, 6, 2, 3)]


@pytest.mark.parametrize("n,bs,hw,ntau", SHAPES)
def test_device_operators_match_oracle(n, bs, hw, ntau):
    p = BandedTauPencil.generate(n, bs, hw, ntau, seed=11)
    sigma = 0.5 + 0.25j
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for pc, opc in (("exact", 1), ("block_jacobi", 2)):
        dev = DeviceOperator(p, sigma, pc=pc)
        cpu = orc.CpuPencilOps(p, sigma, opc)
        ya, yb, yp = dev.matvec(x), dev.matvec_B(x), dev.solve(x)
        ra, rb, rp = cpu.apply("orc_matvec_A", x), cpu.apply("orc_matvec_B", x), cpu.apply("orc_precond", x)
        assert np.linalg.norm(ya - ra) <= 1e-13 * np.linalg.norm(ra)
        assert np.linalg.norm(yb - rb) <= 1e-15 * np.linalg.norm(rb) + 1e-300
        # the preconditioner solves M y = x: both are exact solves, compare the
        # solutions and the residual of the device one
        assert np.linalg.norm(yp - rp) <= 1e-9 * np.linalg.norm(rp), pc
        if pc == "exact":
            # two right-hand sides in one pass: each as the one-vector solve
            x2 = rng.standard_normal(n) - 0.5j * rng.standard_normal(n)
            z0, z1 = dev.solve2(x, x2)
            assert np.linalg.norm(z0 - yp) <= 1e-14 * np.linalg.norm(yp)
            r2 = cpu.apply("orc_precond", x2)
            assert np.linalg.norm(z1 - r2) <= 1e-9 * np.linalg.norm(r2)
        dev.free()


@pytest.mark.parametrize("n,bs,hw,ntau", SHAPES + [(20000, 8, 1, 4), (40000, 8, 2, 16), (3000, 4, 1, 20)])
def test_fused_matvec_ab_matches_separate(n, bs, hw, ntau):
    """bbjd_op_matvec_AB (A x and B x in one pass over x): the band rows and
    B x bit-identical to the separate matvecs, A x to 1e-12 of scipy's
    (the dense rows' dots sum in another order); the complex64 form to FP32
    rounding; more than 16 dense rows take the separate kernels"""
    p = BandedTauPencil.generate(n, bs, hw, ntau, seed=4)
    dev = DeviceOperator(p, 0.2 + 0.1j, pc="none", c64=True)
    rng = np.random.default_rng(n + ntau)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    A, B = p.to_scipy()
    ra_ref, rb_ref = A @ x, B @ x
    try:
        ya, yb = dev.matvec_AB(x)
        ra, rb = dev.matvec(x), dev.matvec_B(x)
        n1 = n - ntau
        np.testing.assert_array_equal(ya[:n1], ra[:n1])
        np.testing.assert_array_equal(yb, rb)
        if ntau > 16:
            np.testing.assert_array_equal(ya, ra)
        assert np.linalg.norm(ya - ra_ref) <= 1e-12 * np.linalg.norm(ra_ref)
        assert np.linalg.norm(yb - rb_ref) <= 1e-15 * np.linalg.norm(rb_ref) + 1e-300
        sa, sb = dev.matvec_AB(x, single=True)
        assert np.linalg.norm(sa - ra_ref) <= 1e-5 * np.linalg.norm(ra_ref)
        assert np.linalg.norm(sb - rb_ref) <= 1e-6 * np.linalg.norm(rb_ref) + 1e-300
        # with the residual of the convergence test, r = tb A x - ta B x and ||r||
        tb, ta = 0.8 - 0.3j, 1.7 + 0.4j
        qa, qb, r, rn = dev.matvec_ABr(x, tb, -ta)
        np.testing.assert_array_equal(qa[:n1], ya[:n1])
        np.testing.assert_array_equal(qb, yb)
        assert np.linalg.norm(qa - ra_ref) <= 1e-12 * np.linalg.norm(ra_ref)
        r_ref = tb * ra_ref - ta * rb_ref
        assert np.linalg.norm(r - r_ref) <= 1e-12 * np.linalg.norm(r_ref)
        assert abs(rn - np.linalg.norm(r_ref)) <= 1e-12 * np.linalg.norm(r_ref)
    finally:
        dev.free()


@pytest.mark.parametrize("n,bs,hw,ntau", [(60000, 8, 1, 4), (40000, 8, 2, 16), (30004, 4, 1, 2), (200000, 8, 1, 4)])
def test_bcr_level_groups_match_level_by_level(n, bs, hw, ntau):
    """the exact solve with 2, 3 or 4 block-cyclic-reduction levels per
    launch (bbjd_set_bcr_group; super-blocks of 8, 16 and 4 rows, the last
    one ragged) equals the one-launch-per-level solve exactly -- the same
    row arithmetic in the same order -- for one and two right-hand sides,
    and solves M y = x as the oracle's banded LU does; with the Schur
    border's dots formed in the last backward launch
    (bbjd_set_bcr_border_fuse; another summation order) to 1e-13.  The
    grouped launches that load every phase's factors at launch
    (bbjd_set_bcr_prefetch, s = 4 and 8) equal the phase-by-phase ones
    exactly, with and without the fused dots"""
    lib = _lib.load()
    p = BandedTauPencil.generate(n, bs, hw, ntau, seed=21)
    sigma = 0.3 - 0.2j
    dev = DeviceOperator(p, sigma, pc="exact")
    rng = np.random.default_rng(8)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x2 = rng.standard_normal(n) - 1j * rng.standard_normal(n)
    try:
        lib.bbjd_set_bcr_border_fuse(0)
        lib.bbjd_set_bcr_group(1)
        y1 = dev.solve(x)
        z0, z1 = dev.solve2(x, x2)
        for g in (2, 3, 4):
            lib.bbjd_set_bcr_group(g)
            for pf in (0, 1):
                lib.bbjd_set_bcr_prefetch(pf)
                np.testing.assert_array_equal(dev.solve(x), y1)
                a, b = dev.solve2(x, x2)
                np.testing.assert_array_equal(a, z0)
                np.testing.assert_array_equal(b, z1)
        lib.bbjd_set_bcr_border_fuse(1)
        for g in (2, 4):
            lib.bbjd_set_bcr_group(g)
            lib.bbjd_set_bcr_prefetch(0)
            f1 = dev.solve(x)
            f2 = dev.solve2(x, x2)
            lib.bbjd_set_bcr_prefetch(1)
            np.testing.assert_array_equal(dev.solve(x), f1)
            for u, v in zip(dev.solve2(x, x2), f2):
                np.testing.assert_array_equal(u, v)
            assert np.linalg.norm(f1 - y1) <= 1e-13 * np.linalg.norm(y1)
            assert np.linalg.norm(f2[0] - z0) <= 1e-13 * np.linalg.norm(z0)
            assert np.linalg.norm(f2[1] - z1) <= 1e-13 * np.linalg.norm(z1)
    finally:
        lib.bbjd_set_bcr_group(4)
        lib.bbjd_set_bcr_border_fuse(1)
        lib.bbjd_set_bcr_prefetch(1)
        dev.free()
    cpu = orc.CpuPencilOps(p, sigma, 1)
    r1, r2 = cpu.apply("orc_precond", x), cpu.apply("orc_precond", x2)
    assert np.linalg.norm(y1 - r1) <= 1e-9 * np.linalg.norm(r1)
    assert np.linalg.norm(z1 - r2) <= 1e-9 * np.linalg.norm(r2)


@pytest.mark.parametrize("n,bs,hw,ntau", [(30004, 4, 1, 2), (60000, 8, 1, 4), (4096, 4, 1, 2), (40000, 8, 2, 16),
                                          (4000, 4, 1, 0)])
def test_c64_preconditioner_direct_io_matches_conversions(n, bs, hw, ntau):
    """bbjd_op_precond_c64 with the exact preconditioner reading and writing
    complex64 inside the solve's own launches (bbjd_set_c64_direct_io; s = 4
    and 8 with dense rows) equals the widen / FP64 solve / round sequence
    exactly, and that equals the FP64 solve of the widened vector rounded to
    complex64; s = 16 and no dense rows take the conversions"""
    lib = _lib.load()
    p = BandedTauPencil.generate(n, bs, hw, ntau, seed=31)
    sigma = 0.4 + 0.3j
    dev = DeviceOperator(p, sigma, pc="exact", c64=True)
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    try:
        lib.bbjd_set_c64_direct_io(0)
        a = dev.solve_c64(x)
        lib.bbjd_set_c64_direct_io(1)
        b = dev.solve_c64(x)
        y = dev.solve(x.astype(np.complex128))
    finally:
        lib.bbjd_set_c64_direct_io(1)
        dev.free()
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(b, y.astype(np.complex64))


def _cfg1(n=2000, k=8, seed=1):
    return BandedTauPencil.generate(n, 4, 1, 2, seed=seed), 0.5 + 0.5j


def test_config1_device_operators_match_oracle():
    p, sigma = _cfg1()
    ops = orc.CpuPencilOps(p, sigma, 1)
    ref = orc.solve("oracle", p.n, 8, sigma, tol=1e-10, maxsub=40, minsub=20, maxiter=500, ops=ops)
    dev = DeviceOperator(p, sigma, pc="exact")
    ev, X, info = jdqz(dev, k=8, sigma=sigma, tol=1e-10, maxsub=40, minsub=20, maxiter=500,
                       preconditioner=dev, return_info=True)
    assert info["nconv"] == 8 == ref["nconv"]
    assert rel_match(ev, ref["evals"]) <= 1e-8
    assert rel_match(ref["evals"], ev) <= 1e-8
    A, B = p.to_scipy()
    for j in range(8):
        th = ev[j]
        r = np.linalg.norm(A @ X[:, j] - th * (B @ X[:, j]))
        assert r <= 1e-8
    assert np.all(info["residuals"] < 1e-10)
    # trajectory: same algorithm, different rounding -> iteration counts close
    assert abs(info["iters"] - ref["iters"]) <= max(5, ref["iters"] // 10)


def test_host_callbacks_diagonal():
    """reference tests/test_bbjd.py:162-174 through the host-callback path"""
    from scipy.sparse.linalg import splu
    A, B = make_diagonal(50)
    lu = splu(A.tocsc())
    ev, _, info = jdqz(A, B, k=5, sigma=0.0, tol=1e-10, preconditioner=lu, return_info=True)
    assert info["nconv"] == 5
    assert np.allclose(np.sort(ev.real), [0.1, 0.2, 0.3, 0.4, 0.5], atol=1e-10)


def test_host_callbacks_match_oracle_trajectory():
    """same operators in host space: the engine follows the oracle's steps"""
    p, sigma = _cfg1(n=400)
    ops = orc.CpuPencilOps(p, sigma, 1)
    ref = orc.solve("oracle", p.n, 4, sigma, tol=1e-10, maxsub=30, minsub=12, maxiter=400, ops=ops)
    out = jdqz_c(p.n, 4, sigma, 0, lambda x: ops.apply("orc_matvec_A", x),
                 lambda x: ops.apply("orc_matvec_B", x), lambda x: ops.apply("orc_precond", x),
                 tol=1e-10, maxsub=30, minsub=12, maxiter=400)
    ev, _, res, nconv, iters, status = out
    assert nconv == ref["nconv"] == 4
    assert rel_match(ev, ref["evals"]) <= 1e-8
    assert abs(iters - ref["iters"]) <= 3


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_inner_gmres_approximate_preconditioner(precision):
    """inner GMRES(20) with an approximate preconditioner -- the exact LU at a
    shift 0.5 away, as in reference tests/test_bbjd.py:459-479 -- in FP64 or
    complex64: same eigenvalues as the oracle's FP64 GMRES, in no more outer
    iterations than the single-step correction needs."""
    p = BandedTauPencil.generate(3200, 8, 2, 8, seed=5)
    sigma = 1 + 2j
    ops = orc.CpuPencilOps(p, sigma + 0.5, 1)
    ref = orc.solve("oracle", p.n, 4, sigma, tol=1e-10, maxsub=40, minsub=16, maxiter=400,
                    inner_maxiter=20, inner_tol=1e-3, ops=ops)
    assert ref["nconv"] == 4
    dev = DeviceOperator(p, sigma + 0.5, pc="exact", c64=True)
    ev, X, info = jdqz(dev, k=4, sigma=sigma, tol=1e-10, maxsub=40, minsub=16, maxiter=400,
                       inner_maxiter=20, inner_tol=1e-3, preconditioner=dev, return_info=True,
                       precision=precision)
    assert info["nconv"] == 4
    assert rel_match(ev, ref["evals"]) <= 1e-8
    assert rel_match(ref["evals"], ev) <= 1e-8
    A, B = p.to_scipy()
    for j in range(4):
        assert np.linalg.norm(A @ X[:, j] - ev[j] * (B @ X[:, j])) <= 1e-8
    ev0, _, info0 = jdqz(dev, k=4, sigma=sigma, tol=1e-10, maxsub=40, minsub=16, maxiter=400,
                         preconditioner=dev, return_info=True)
    assert info["iters"] <= info0["iters"]


def test_inner_projected_preconditioner_matches_oracle():
    """complex64 inner GMRES(5) in the projected-preconditioner form
    (inner_form 'projected': one oblique projection (I - Yh H^-1 Q^H) per
    Arnoldi step instead of two projections on each side) with the exact LU,
    on the serving workload's shape (block size 8, block-tridiagonal, 4 tau
    rows): the oracle's eigenvalue set (FP64, one-step correction) within
    1e-8 both ways, every pair's FP64 residual from a fresh matvec below tol,
    the form taken by every inner solve, and about as many outer iterations
    as the reference form's inner solve with the same budget"""
    from bbjd_amd.configs import row_norm_bound
    p = BandedTauPencil.generate(20000, 8, 1, 4, seed=31)
    sigma = 1.5 - 2.2j
    ref = orc.solve("oracle", p.n, 8, sigma, tol=1e-10, maxsub=32, minsub=16, maxiter=1000,
                    ops=orc.CpuPencilOps(p, sigma, 1))
    assert ref["nconv"] == 8
    dev = DeviceOperator(p, sigma, pc="exact", c64=True)
    kw = dict(tol=1e-10, maxsub=32, minsub=16, maxiter=1000, inner_maxiter=5, inner_tol=1e-2, device_op=dev,
              precision="mixed", return_stats=True)
    ev, X, res, nconv, iters, status, st = jdqz_c(p.n, 8, sigma, 0, None, None, True, inner_form="projected", **kw)
    rf = jdqz_c(p.n, 8, sigma, 0, None, None, True, inner_form="reference", **kw)
    print(f"[projected] iters {iters} (reference form {rf[4]}, oracle plain {ref['iters']}), inner steps "
          f"{st['inner_steps']} (reference form {rf[6]['inner_steps']}), projected calls {st['inner_projected']}",
          flush=True)
    assert status == 0 and nconv == 8
    assert st["inner_projected"] == st["outer_iters"] - st["lock_iters"] > 0
    assert st["inner_projected_fallbacks"] == 0
    assert rf[6]["inner_projected"] == 0
    assert rel_match(ev, ref["evals"]) <= 1e-8 and rel_match(ref["evals"], ev) <= 1e-8
    A, B = p.to_scipy()
    anorm = row_norm_bound(p)
    for j in range(8):
        x = X[:, j]
        assert np.linalg.norm(A @ x - ev[j] * (B @ x)) / ((anorm + abs(ev[j])) * np.linalg.norm(x)) <= 1e-10
    # the same budget of inner steps converges alike in both forms (64 and
    # 64 outer iterations here); the projected form's steps are cheaper
    assert iters <= 1.2 * rf[4]


def test_serving_loop_projected_inner_alone_equal_mix():
    """the served form of the serving workload: complex64 inner GMRES(5) in
    the projected-preconditioner form with the exact LU, block size 8,
    4 tau rows; each request's eigenvalues bit-identical alone and in a mix
    of four workers, and each request's set within 1e-8 of the oracle's
    plain FP64 solve both ways"""
    from bbjd_amd import serve as srv
    rng = np.random.default_rng(17)
    jobs, pens = [], []
    for i, (n, k) in enumerate(((20000, 8), (20000, 16), (24000, 8), (20000, 8), (16000, 16))):
        p = BandedTauPencil.generate(n, 8, 1, 4, seed=300 + i)
        s = complex(rng.uniform(-4, 4), rng.uniform(-4, 4))
        pens.append(p)
        jobs.append(dict(pencil=p, pc="exact", sigma=s, nev=k, maxsub=4 * k, minsub=2 * k, maxiter=1000,
                         inner_maxiter=5, inner_tol=1e-2, tol=1e-10))
    kw = dict(realtime=False, precision="mixed", inner_form="projected")
    mix, rep = srv.serve(jobs, workers=4, **kw)
    assert rep["failed"] == 0 and rep["peak_concurrency"] >= 2
    for i, j in enumerate(jobs):
        alone, _ = srv.serve([j], workers=1, **kw)
        a = alone[0]
        assert a["status"] == 0 and a["nconv"] == j["nev"], i
        np.testing.assert_array_equal(mix[i]["eigenvalues"], a["eigenvalues"], err_msg=str(i))
        ref = orc.solve("oracle", pens[i].n, j["nev"], j["sigma"], tol=1e-10, maxsub=j["maxsub"],
                        minsub=j["minsub"], maxiter=1000, ops=orc.CpuPencilOps(pens[i], j["sigma"], 1))
        assert ref["nconv"] == j["nev"], i
        assert rel_match(a["eigenvalues"], ref["evals"]) <= 1e-8, i
        assert rel_match(ref["evals"], a["eigenvalues"]) <= 1e-8, i


@pytest.mark.parametrize("precision,pc", [("mixed", "exact"), ("fp64", "exact"), ("mixed", "block_jacobi")])
def test_inner_gmres_graphs_replay_identically(precision, pc):
    """the inner GMRES prologue and Arnoldi steps captured into HIP graphs
    (bbjd_gpu_options.graphs) and replayed on later outer iterations run the
    eager path's kernels: identical eigenvalues, vectors, residuals, outer
    iterations and inner steps; segments are captured once per deflation
    count and replayed after that"""
    p = BandedTauPencil.generate(6000, 4, 1, 2, seed=11)
    sigma = 0.4 + 0.3j
    dev = DeviceOperator(p, sigma + (0.5 if pc == "exact" else 0.0), pc=pc, c64=True)
    kw = dict(tol=1e-10, maxsub=32, minsub=16, maxiter=300, inner_maxiter=20, inner_tol=1e-3,
              device_op=dev, precision=precision, return_stats=True)
    eager = jdqz_c(p.n, 6, sigma, 0, None, None, True, graphs=False, **kw)
    graph = jdqz_c(p.n, 6, sigma, 0, None, None, True, graphs=True, **kw)
    se, sg = eager[6], graph[6]
    print(f"[graphs] {precision}/{pc}: nconv {eager[3]} iters {eager[4]} inner {se['inner_steps']} "
          f"captures {sg['graph_captures']} launches {sg['graph_launches']}", flush=True)
    assert se["graph_captures"] == 0 and se["graph_launches"] == 0
    assert eager[3] == graph[3] and eager[4] == graph[4] and eager[5] == graph[5]
    assert se["inner_steps"] == sg["inner_steps"] > 0
    np.testing.assert_array_equal(eager[0], graph[0])
    np.testing.assert_array_equal(eager[1], graph[1])
    np.testing.assert_array_equal(eager[2], graph[2])
    assert 0 < sg["graph_captures"] < sg["graph_launches"]
    # one launch per prologue and per Arnoldi step
    outer_with_gmres = sg["outer_iters"] - sg["lock_iters"]
    assert sg["graph_launches"] == sg["inner_steps"] + outer_with_gmres
    if pc == "exact":
        assert eager[3] == 6
        A, B = p.to_scipy()
        X = graph[1]
        for j in range(6):
            assert np.linalg.norm(A @ X[:, j] - graph[0][j] * (B @ X[:, j])) <= 1e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_inner_gmres_lookahead_matches_one_step_at_a_time(precision):
    """the inner GMRES with 2, 4 and 8 Arnoldi steps in flight
    (bbjd_set_gmres_lookahead) against one step at a time: identical
    eigenvalues, vectors, residuals, outer iterations and inner steps used;
    steps run past the stop are counted and discarded"""
    lib = _lib.load()
    p = BandedTauPencil.generate(5000, 4, 1, 2, seed=13)
    sigma = 0.2 - 0.5j
    dev = DeviceOperator(p, sigma + 0.4, pc="exact", c64=True)
    kw = dict(tol=1e-10, maxsub=32, minsub=16, maxiter=300, inner_maxiter=20, inner_tol=1e-3,
              device_op=dev, precision=precision, return_stats=True)
    runs = {}
    try:
        for look in (1, 2, 4, 8):
            lib.bbjd_set_gmres_lookahead(look)
            runs[look] = jdqz_c(p.n, 6, sigma, 0, None, None, True, **kw)
    finally:
        lib.bbjd_set_gmres_lookahead(4)
    one = runs[1]
    print(f"[lookahead] {precision}: nconv {one[3]} iters {one[4]} inner {one[6]['inner_steps']}; discarded "
          f"{[runs[k][6]['inner_steps_discarded'] for k in (1, 2, 4, 8)]}", flush=True)
    assert one[3] == 6 and one[6]["inner_steps"] > 0 and one[6]["inner_steps_discarded"] == 0
    for look in (2, 4, 8):
        r = runs[look]
        assert r[3] == one[3] and r[4] == one[4] and r[5] == one[5]
        assert r[6]["inner_steps"] == one[6]["inner_steps"]
        np.testing.assert_array_equal(r[0], one[0])
        np.testing.assert_array_equal(r[1], one[1])
        np.testing.assert_array_equal(r[2], one[2])
    assert 0 < runs[2][6]["inner_steps_discarded"] <= runs[4][6]["inner_steps_discarded"] \
        <= runs[8][6]["inner_steps_discarded"]


def test_correction_gmres_entry_matches_oracle():
    """bbjd_correction_gmres (reference include/bbjd.h:268-282 with the
    argument order of src/jd_core.c:568-582) on host arrays vs the oracle's
    GMRES on the same projected correction equation"""
    lib = _lib.load()
    p = BandedTauPencil.generate(1200, 4, 1, 2, seed=9)
    sigma = 0.3 + 0.1j
    ops = orc.CpuPencilOps(p, sigma + 0.3, 1)
    n = p.n
    rng = np.random.default_rng(1)
    u = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    u /= np.linalg.norm(u)
    w = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    w /= np.linalg.norm(w)
    r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ta, tb = 0.7 - 0.2j, 0.9 + 0.1j
    z_ref = orc.gmres(n, u[:, None], w[:, None], ta, tb, r, ops, maxiter=15, tol=1e-6)
    prm = _lib.Params()
    lib.bbjd_params_init(C.byref(prm))
    prm.n, prm.nev, prm.inner_maxiter, prm.maxsub = n, 2, 15, 20
    ws = lib.bbjd_workspace_alloc(C.byref(prm))
    assert ws
    z = np.zeros(n, np.complex128)
    cbs = [_lib.CALLBACK(lambda nn, x, y, d, f=f: _apply_into(ops, f, nn, x, y))
           for f in ("orc_matvec_A", "orc_matvec_B", "orc_precond")]
    rc = lib.bbjd_correction_gmres(n, u.ctypes.data, w.ctypes.data, _lib.zc(ta), _lib.zc(tb),
                                   r.ctypes.data, z.ctypes.data,
                                   *[C.cast(c, C.c_void_p).value for c in cbs], None, 15, 1e-6, ws)
    lib.bbjd_workspace_free(ws)
    assert rc == 0
    assert np.linalg.norm(z - z_ref) <= 1e-9 * np.linalg.norm(z_ref)
    assert abs(np.vdot(u, z)) <= 1e-12 * np.linalg.norm(z)


def _apply_into(ops, name, n, x, y):
    xa = np.frombuffer((C.c_double * (2 * n)).from_address(x), dtype=np.complex128)
    ya = np.frombuffer((C.c_double * (2 * n)).from_address(y), dtype=np.complex128)
    ya[:] = ops.apply(name, xa.copy())
    return 0


def test_thick_restart_small_maxsub():
    """forces several restarts (maxsub 16): parity with the oracle"""
    p, sigma = _cfg1(n=800, seed=3)
    ops = orc.CpuPencilOps(p, sigma, 1)
    ref = orc.solve("oracle", p.n, 6, sigma, tol=1e-10, maxsub=16, minsub=8, maxiter=600, ops=ops)
    dev = DeviceOperator(p, sigma, pc="exact")
    ev, _, info = jdqz(dev, k=6, sigma=sigma, tol=1e-10, maxsub=16, minsub=8, maxiter=600,
                       preconditioner=dev, return_info=True)
    assert info["nconv"] == ref["nconv"] == 6
    assert rel_match(ev, ref["evals"]) <= 1e-8


@pytest.mark.parametrize("space", ["device", "host"])
def test_search_space_exhaustion_breakdown(space):
    """tol below rounding: the search space fills C^n, the next expansion and
    its random restart vector both vanish under CGS2 and the solve ends with
    BBJD_ERR_BREAKDOWN at the same iteration as the oracle (reference
    src/jd_core.c:1032-1058).  The device path defers the breakdown test to
    the pencil update and recomputes the column along this branch."""
    for n, k, ms, mn, nt in ((16, 6, 30, 14, 2), (8, 3, 12, 4, 0)):
        p = BandedTauPencil.generate(n, 4, 1, nt, seed=3)
        sigma = 0.3 + 0.1j
        ops = orc.CpuPencilOps(p, sigma, 1)
        ref = orc.solve("oracle", n, k, sigma, tol=1e-17, maxsub=ms, minsub=mn, maxiter=300, ops=ops)
        assert ref["status"] == -5 and ref["nconv"] == 0
        if space == "device":
            dev = DeviceOperator(p, sigma, pc="exact")
            out = jdqz_c(n, k, sigma, 0, None, None, dev, tol=1e-17, maxsub=ms, minsub=mn, maxiter=300,
                         device_op=dev)
        else:
            out = jdqz_c(n, k, sigma, 0, lambda x: ops.apply("orc_matvec_A", x),
                         lambda x: ops.apply("orc_matvec_B", x), lambda x: ops.apply("orc_precond", x),
                         tol=1e-17, maxsub=ms, minsub=mn, maxiter=300)
        ev, _, res, nconv, iters, status = out
        assert status == -5 and nconv == 0, (n, status, nconv)
        assert iters == ref["iters"], (n, iters, ref["iters"])


@pytest.mark.parametrize("n,p,q", [(5000, 64, 64), (70001, 17, 40), (300, 130, 7), (33, 1, 1)])
def test_block_gram_matches_numpy(n, p, q):
    """the restart's device Gram X^H Y (ragged tiles, row chunks) against
    numpy; deterministic across calls"""
    lib = _lib.load()
    rng = np.random.default_rng(n + p + q)
    X = np.asfortranarray(rng.standard_normal((n, p)) + 1j * rng.standard_normal((n, p)))
    Y = np.asfortranarray(rng.standard_normal((n, q)) + 1j * rng.standard_normal((n, q)))
    G = np.zeros((p, q), np.complex128, order="F")
    assert lib.bbjd_block_gram(n, p, X.ctypes.data, q, Y.ctypes.data, G.ctypes.data) == 0
    ref = X.conj().T @ Y
    scale = np.linalg.norm(X, axis=0)[:, None] * np.linalg.norm(Y, axis=0)[None, :]
    assert np.max(np.abs(G - ref) / scale) <= 1e-14
    G2 = np.zeros_like(G)
    assert lib.bbjd_block_gram(n, p, X.ctypes.data, q, Y.ctypes.data, G2.ctypes.data) == 0
    assert np.array_equal(G, G2)


@pytest.mark.parametrize("n,k,keep", [(5000, 64, 32), (70001, 128, 64), (3001, 200, 80), (1000, 9, 3),
                                       (50, 300, 250), (20011, 256, 128), (4099, 100, 96), (777, 17, 17)])
def test_block_rotate_matches_numpy(n, k, keep):
    """the restart's in-place rotation X[:, :keep] = X Z against numpy: the
    streaming matrix-core form (mode 1: input columns through LDS in chunks,
    every output width up to 128 and ragged row tiles), the staged
    matrix-core form (mode 2: 32 or 16 rows per block; k = 200 exceeds the
    32-row tile) and the vector kernel (k = 300, and mode 0)"""
    lib = _lib.load()
    rng = np.random.default_rng(n + k + keep)
    X = np.asfortranarray(rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k)))
    Z = np.asfortranarray(rng.standard_normal((k, keep)) + 1j * rng.standard_normal((k, keep)))
    ref = X @ Z
    try:
        for mfma in (1, 2, 0):
            lib.bbjd_set_rotate_mfma(mfma)
            Xg = X.copy(order="F")
            assert lib.bbjd_block_rotate(n, k, keep, Xg.ctypes.data, Z.ctypes.data) == 0
            err = np.linalg.norm(Xg[:, :keep] - ref) / np.linalg.norm(ref)
            assert err <= 1e-14, (mfma, err)
            assert np.array_equal(Xg[:, keep:], X[:, keep:])        # columns past keep untouched
    finally:
        lib.bbjd_set_rotate_mfma(1)


def test_qz_modes_same_eigenpairs():
    """the eigenvalue-only projected step with inverse-iteration Ritz vectors
    ('ritz') and the ordered Schur form every iteration ('schur', the
    reference's LAPACK sequence) converge to the same eigenvalues, both equal
    to the oracle's; restarts included"""
    p, sigma = _cfg1(n=2000, seed=1)
    ops = orc.CpuPencilOps(p, sigma, 1)
    ref = orc.solve("oracle", p.n, 8, sigma, tol=1e-10, maxsub=20, minsub=10, maxiter=600, ops=ops)
    dev = DeviceOperator(p, sigma, pc="exact")
    out = {}
    for mode in ("ritz", "schur"):
        ev, X, res, nconv, iters, status, st = jdqz_c(p.n, 8, sigma, 0, None, None, dev, tol=1e-10, maxsub=20,
                                                       minsub=10, maxiter=600, device_op=dev, qz_mode=mode,
                                                       return_stats=True)
        assert status == 0 and nconv == 8 == ref["nconv"], mode
        assert st["restarts"] > 0
        assert rel_match(ev, ref["evals"]) <= 1e-8 and rel_match(ref["evals"], ev) <= 1e-8, mode
        assert np.all(res < 1e-10)
        out[mode] = (ev, iters)
    assert rel_match(out["ritz"][0], out["schur"][0]) <= 1e-8
    assert abs(out["ritz"][1] - out["schur"][1]) <= max(5, ref["iters"] // 10)


def _serve_jobs(seed=7):
    rng = np.random.default_rng(seed)
    shapes = [(2000, 4), (5000, 8), (12000, 4), (3000, 8), (8000, 6), (2000, 8), (6000, 4), (4000, 8)]
    jobs = []
    for i, (n, k) in enumerate(shapes):
        p = BandedTauPencil.generate(n, 8, 1, 4, seed=100 + i)
        sigma = complex(rng.uniform(-3, 3), rng.uniform(-3, 3))
        jobs.append(dict(pencil=p, pc="exact", sigma=sigma, nev=k, maxsub=4 * k, minsub=2 * k, maxiter=1000,
                         tol=1e-10, arrival=0.02 * i))
    return jobs


def test_serving_loop_results_alone_equal_mix():
    """bbjd_serve (continuous batching over resident workspaces, operators
    uploaded and factored on admission): each request's eigenvalues are
    bit-identical alone and in a concurrent mix -- arrivals in real time,
    largest-first admission, windowed admission, and an HBM budget that
    throttles admission -- every pair's FP64 residual ||A x - lambda B x|| is
    below tol (||A|| + |lambda| ||B||), and two requests match the oracle.
    With one worker and every request waiting from the start, the windowed
    order admits the largest (nev, then n) of the 3 longest-waiting requests
    each time, the oldest on a tie"""
    from bbjd_amd import serve as srv
    from bbjd_amd.configs import row_norm_bound
    jobs = _serve_jobs()
    mix, rep = srv.serve(jobs, workers=4, realtime=True)
    assert rep["failed"] == 0 and rep["peak_concurrency"] >= 2
    lpt, _ = srv.serve(jobs, workers=4, realtime=False, order="largest")
    win, _ = srv.serve(jobs, workers=1, realtime=False, order="window", window=3)
    pend, expect = list(range(len(jobs))), []
    while pend:
        best = pend[0]
        for c in pend[1:3]:
            if (jobs[c]["nev"], jobs[c]["pencil"].n) > (jobs[best]["nev"], jobs[best]["pencil"].n):
                best = c
        expect.append(best)
        pend.remove(best)
    assert sorted(range(len(jobs)), key=lambda i: win[i]["t_admit"]) == expect
    budget = rep["peak_bytes"] // 2
    tight, rep_t = srv.serve(jobs, workers=4, realtime=False, hbm_budget=budget)
    # admission keeps within the budget, except that a request runs alone
    # when it does not fit beside the others
    ops = [DeviceOperator(j["pencil"], j["sigma"], pc="exact") for j in jobs]
    largest_op = max(o.device_bytes for o in ops)
    for o in ops:
        o.free()
    assert rep_t["failed"] == 0 and rep_t["peak_concurrency"] < rep["peak_concurrency"] + 1
    assert rep_t["peak_bytes"] <= max(budget, rep_t["workers"] * rep_t["workspace_bytes"] + largest_op)
    print(f"[serve] mix peak {rep['peak_bytes']} concurrency {rep['peak_concurrency']}; tight budget {budget}: "
          f"workers {rep_t['workers']} peak {rep_t['peak_bytes']} concurrency {rep_t['peak_concurrency']}",
          flush=True)
    for i, j in enumerate(jobs):
        alone, _ = srv.serve([j], workers=1, realtime=False)
        a = alone[0]
        assert a["status"] == 0 and a["nconv"] == j["nev"], i
        for tag, other in (("mix", mix[i]), ("largest-first", lpt[i]), ("window", win[i]), ("tight", tight[i])):
            np.testing.assert_array_equal(other["eigenvalues"], a["eigenvalues"], err_msg=f"{i} {tag}")
        anorm = row_norm_bound(j["pencil"])           # lower bound of ||A||: a stricter test
        assert np.all(a["pair_residuals"] <= 1e-10 * (anorm + np.abs(a["eigenvalues"]))), i
    for i in (0, 5):
        j = jobs[i]
        p = j["pencil"]
        ref = orc.solve("oracle", p.n, j["nev"], j["sigma"], tol=1e-10, maxsub=j["maxsub"], minsub=j["minsub"],
                        maxiter=1000, ops=orc.CpuPencilOps(p, j["sigma"], 1))
        assert ref["nconv"] == j["nev"]
        assert rel_match(mix[i]["eigenvalues"], ref["evals"]) <= 1e-8
        assert rel_match(ref["evals"], mix[i]["eigenvalues"]) <= 1e-8


def test_resident_server_runs_batches_without_allocating():
    """bbjd_server_create / _run / _destroy (serve.Server): the same batch run
    twice on one resident server, once with operators built on admission
    (host pencils) and once with resident device operators, gives
    eigenvalues bit-identical to the one-shot loop; free device memory does
    not drop across runs (no workspace allocated per run); the
    projected step on the device (qz_mode "device") gives the same
    eigenvalues to 1e-8; a closed server refuses to run"""
    from bbjd_amd import serve as srv
    jobs = _serve_jobs(seed=11)[:6]
    once, _ = srv.serve(jobs, workers=3, realtime=False)
    ops = [DeviceOperator(j["pencil"], j["sigma"], pc="exact") for j in jobs]
    resident = [dict(j, op=o) for j, o in zip(jobs, ops)]
    for r in resident:
        r.pop("pencil")
    try:
        with srv.Server(jobs, workers=3) as server:
            assert server.workers == 3 and server.workspace_bytes > 0
            free0 = srv.device_free_bytes()
            a, rep_a = server.run(jobs, realtime=False)
            free1 = srv.device_free_bytes()
            b, rep_b = server.run(resident, realtime=False)
            free2 = srv.device_free_bytes()
            print(f"[server] free {free0} -> {free1} -> {free2}; peaks {rep_a['peak_bytes']} {rep_b['peak_bytes']}",
                  flush=True)
            # a run allocates no workspace (each is tens of MB here); on-demand
            # operators are freed when their request ends
            assert free1 >= free0 - (8 << 20) and free2 >= free1 - (8 << 20)
            assert rep_a["failed"] == rep_b["failed"] == 0
            assert rep_b["peak_bytes"] <= rep_a["peak_bytes"]        # nothing to build on admission
            for i in range(len(jobs)):
                np.testing.assert_array_equal(a[i]["eigenvalues"], once[i]["eigenvalues"], err_msg=str(i))
                np.testing.assert_array_equal(b[i]["eigenvalues"], once[i]["eigenvalues"], err_msg=str(i))
        with pytest.raises(RuntimeError):
            server.run(jobs)
        dev, _ = srv.serve(jobs, workers=3, realtime=False, qz_mode="device")
        for i in range(len(jobs)):
            assert dev[i]["nconv"] == once[i]["nconv"]
            assert rel_match(dev[i]["eigenvalues"], once[i]["eigenvalues"]) <= 1e-8
    finally:
        for o in ops:
            o.free()


def test_serving_loop_arrival_waves():
    """arrival waves (bbjd_request.wave): the same six requests submitted as
    three waves of one run on a resident server with two workers; a wave's
    arrivals count from the admission of the previous wave's last request
    (t_arrive = that admission + arrival), no request is admitted before it
    arrives or before every request of an earlier wave, and each copy's
    eigenvalues are bit-identical to the single-wave run"""
    from bbjd_amd import serve as srv
    jobs = _serve_jobs(seed=13)[:6]
    ops = [DeviceOperator(j["pencil"], j["sigma"], pc="exact") for j in jobs]
    resident = [dict(j, op=o) for j, o in zip(jobs, ops)]
    for r in resident:
        r.pop("pencil")
    try:
        with srv.Server(resident, workers=2) as server:
            once, _ = server.run(resident, realtime=True)
            waves = [dict(r, wave=w) for w in range(3) for r in resident]
            res, rep = server.run(waves, realtime=True)
        assert rep["failed"] == 0
        m = len(resident)
        for w in range(3):
            part = res[w * m:(w + 1) * m]
            base = 0.0 if w == 0 else max(r["t_admit"] for r in res[(w - 1) * m:w * m])
            for i, r in enumerate(part):
                assert r["wave"] == w
                assert abs(r["t_arrive"] - (base + resident[i]["arrival"])) <= 1e-9, (w, i)
                assert r["t_admit"] >= r["t_arrive"] - 1e-6, (w, i)
                if w > 0:
                    assert r["t_admit"] >= base - 1e-6, (w, i)
                np.testing.assert_array_equal(r["eigenvalues"], once[i]["eigenvalues"], err_msg=f"{w} {i}")
        # with two workers the next wave starts while the previous one still runs
        assert min(r["t_start"] for r in res[m:2 * m]) < max(r["t_done"] for r in res[:m])
    finally:
        for o in ops:
            o.free()


def test_serving_loop_blocking_waits_same_results():
    """workers whose waits sleep until the GPU signals (sync_yield "block":
    bbjd_gpu_options.blocking_sync, events created with
    hipEventBlockingSync) give the same eigenvalues as yielding ones, FP64
    and complex64 inner solves"""
    from bbjd_amd import serve as srv
    jobs = _serve_jobs(seed=17)[:5]
    for prec, inner in (("fp64", 0), ("mixed", 5)):
        js = [dict(j, inner_maxiter=inner) for j in jobs]
        a, _ = srv.serve(js, workers=3, realtime=False, precision=prec)
        b, rep = srv.serve(js, workers=3, realtime=False, precision=prec, sync_yield="block")
        assert rep["failed"] == 0
        for i in range(len(js)):
            np.testing.assert_array_equal(a[i]["eigenvalues"], b[i]["eigenvalues"], err_msg=f"{prec} {i}")
            assert a[i]["iters"] == b[i]["iters"]


def test_serving_loop_mixed_precision_alone_equal_mix():
    """the small-solve form of the serving loop: GMRES(20) in complex64 with
    the exact LU, FP32 accumulation, four workers; each request's eigenvalues
    bit-identical alone and in the mix, every FP64 residual below tol, and
    every eigenvalue of the mixed and of the FP64 inner solve within 1e-8 of
    one of the pencil's 48 eigenvalues nearest sigma, found independently by
    ARPACK on the shift-and-invert operator (A - sigma B)^-1 B (the oracle's
    banded LU).  The two sets are not compared with each other: with inexact
    inner solves JDQZ need not return exactly the nev nearest sigma (one
    request here locks the 20th nearest), and the two inner precisions may
    lock different ones"""
    from scipy.sparse.linalg import LinearOperator, eigs
    from bbjd_amd import serve as srv
    from bbjd_amd.configs import row_norm_bound
    rng = np.random.default_rng(3)
    jobs = []
    for i, n in enumerate((4096, 6000, 9000, 5004, 12000, 7000)):
        p = BandedTauPencil.generate(n, 4, 1, 2, seed=200 + i)
        s = np.sqrt(rng.uniform()) * np.exp(2j * np.pi * rng.uniform())
        jobs.append(dict(pencil=p, pc="exact", sigma=complex(s), nev=8, maxsub=32, minsub=16, maxiter=1000,
                         inner_maxiter=20, inner_tol=1e-3, tol=1e-10))
    mix, rep = srv.serve(jobs, workers=4, realtime=False, precision="mixed")
    f64, _ = srv.serve(jobs, workers=4, realtime=False, precision="fp64")
    assert rep["failed"] == 0 and rep["peak_concurrency"] >= 2
    for i, j in enumerate(jobs):
        alone, _ = srv.serve([j], workers=1, realtime=False, precision="mixed")
        a = alone[0]
        assert a["status"] == 0 and a["nconv"] == 8, i
        np.testing.assert_array_equal(mix[i]["eigenvalues"], a["eigenvalues"], err_msg=str(i))
        anorm = row_norm_bound(j["pencil"])
        assert np.all(a["pair_residuals"] <= 1e-10 * (anorm + np.abs(a["eigenvalues"]))), i
        assert f64[i]["nconv"] == 8
        A, B = j["pencil"].to_scipy()
        lu = orc.CpuPencilOps(j["pencil"], j["sigma"], 1)
        op = LinearOperator(A.shape, matvec=lambda x: lu.apply("orc_precond", B @ x), dtype=np.complex128)
        mu = eigs(op, k=48, which="LM", return_eigenvectors=False, tol=1e-12, ncv=128)
        w = j["sigma"] + 1.0 / mu
        for ev in np.concatenate([a["eigenvalues"], f64[i]["eigenvalues"]]):
            assert np.min(np.abs(w - ev)) <= 1e-8 * max(1.0, abs(ev)), (i, ev)


@pytest.mark.parametrize("pc", ["exact", "block_jacobi"])
def test_serving_loop_grouped_inner_solves_identical(pc):
    """inner solves run in batches on the batcher's threads (inner_batch):
    each Arnoldi step's kernels of the solves in a batch are merged into
    shared launches, each solve's blocks running with its own grid and
    arguments, so every request's eigenvalues, pair residuals and iteration
    count are bit-identical to the workers launching their own kernels.
    Requests of different n (different preconditioner level counts, so
    different launch sequences to merge); the batcher's counters show the
    merging (fewer launches than recorded ones)"""
    from bbjd_amd import serve as srv
    rng = np.random.default_rng(29)
    jobs = []
    for i, n in enumerate((4096, 6000, 9000, 5004, 12000, 7000, 16384, 4500)):
        p = BandedTauPencil.generate(n, 4, 1, 2, seed=500 + i)
        s = np.sqrt(rng.uniform()) * np.exp(2j * np.pi * rng.uniform())
        jobs.append(dict(pencil=p, pc=pc, sigma=complex(s), nev=8 if pc == "exact" else 4, maxsub=32, minsub=16,
                         maxiter=300, inner_maxiter=20, inner_tol=1e-3, tol=1e-10))
    base, rep0 = srv.serve(jobs, workers=4, realtime=False, precision="mixed")
    with srv.Server(jobs, workers=4, precision="mixed", inner_batch=4, inner_batchers=2) as S:
        grp, rep = S.run(jobs, realtime=False)
        st = S.batch_stats()
    assert rep["failed"] == 0 and rep0["failed"] == 0
    assert st["jobs"] > 0 and st["batches"] < st["jobs"] and 0 < st["launches"] < st["recorded"], st
    for i in range(len(jobs)):
        assert grp[i]["status"] == base[i]["status"] and grp[i]["nconv"] == base[i]["nconv"], i
        assert grp[i]["iters"] == base[i]["iters"], i
        np.testing.assert_array_equal(grp[i]["eigenvalues"], base[i]["eigenvalues"], err_msg=str(i))
        np.testing.assert_array_equal(grp[i]["pair_residuals"], base[i]["pair_residuals"], err_msg=str(i))
    if pc == "exact":
        assert all(r["nconv"] == 8 for r in grp)


def test_serving_loop_isolates_failing_requests():
    """a request the engine rejects (maxsub 1: BBJD_ERR_INVALID from the
    host's validation, before any solve kernel) and one that runs out of
    iterations (maxiter 2: status 1, fewer than nev pairs) are reported as
    such, the report counts the one failure, and every other request's
    eigenvalues are bit-identical to those of a run without the two"""
    from bbjd_amd import serve as srv
    good = _serve_jobs()[:4]
    bad = dict(good[1], maxsub=1, arrival=0.01)
    short = dict(good[2], maxiter=2, arrival=0.03)
    res, rep = srv.serve(good[:2] + [bad] + good[2:] + [short], workers=3, realtime=False)
    ref, rep_ref = srv.serve(good, workers=3, realtime=False)
    assert rep_ref["failed"] == 0 and rep["failed"] == 1
    assert res[2]["status"] == -2 and res[2]["nconv"] == 0
    assert res[5]["status"] == 1 and res[5]["nconv"] < short["nev"]
    for a, b in zip(res[:2] + res[3:5], ref):
        assert a["status"] == 0 and a["nconv"] == b["nconv"]
        np.testing.assert_array_equal(a["eigenvalues"], b["eigenvalues"])


def test_svdv_axpy_reduces_dots_bit_identically():
    """bbjd_set_svdv_fused: the Sleijpen-van der Vorst step's axpy launch
    reducing its two dots' partials itself gives the same solve, bit for bit
    (eigenvalues and iteration count), as a finishing launch of its own, at a
    size whose projection grid is too large for the in-kernel finish"""
    lib = _lib.load()
    p = BandedTauPencil.generate(60000, 8, 1, 4, seed=5)
    sigma = 0.2 + 0.7j
    dev = DeviceOperator(p, sigma, pc="exact")
    out = []
    try:
        for f in (0, 1):
            lib.bbjd_set_svdv_fused(f)
            ev, X, info = jdqz(dev, k=4, sigma=sigma, tol=1e-10, maxsub=20, minsub=10, maxiter=300,
                               preconditioner=dev, return_info=True)
            out.append((ev, info["iters"], info["nconv"]))
    finally:
        lib.bbjd_set_svdv_fused(1)
        dev.free()
    assert out[0][2] == out[1][2] == 4
    assert out[0][1] == out[1][1]
    np.testing.assert_array_equal(out[0][0], out[1][0])
