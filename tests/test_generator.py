"""The seeded pencil generator (bbjd_pencil_generate): structure, value
distribution CN(0,1) (+4 on the diagonal of the diagonal blocks), zero band
on tau rows, B = diag(1..1, 0 x ntau), determinism."""
import numpy as np

from bbjd_amd.pencil import BandedTauPencil


def test_structure_and_distribution():
    n, bs, hw, nt = 40000, 8, 2, 16
    p = BandedTauPencil.generate(n, bs, hw, nt, seed=123)
    nb, nd = n // bs, 2 * hw + 1
    blk = p.band.reshape(nb, nd, bs, bs)      # [bi, d, c, r]
    # out-of-range blocks are zero
    assert np.all(blk[0, :hw] == 0) and np.all(blk[-1, hw + 1:] == 0)
    # tau rows: band entries zero; bdiag 1 then 0
    n1 = n - nt
    rows = (np.arange(nb)[:, None] * bs + np.arange(bs)[None, :])          # [bi, r]
    taurow = rows >= n1
    assert np.all(blk.transpose(0, 1, 3, 2)[taurow[:, None, :, None].repeat(nd, 1).repeat(bs, 3)] == 0)
    assert np.all(p.bdiag[:n1] == 1) and np.all(p.bdiag[n1:] == 0)
    # off-diagonal entries: mean 0, E|z|^2 = 1, Re and Im each variance 1/2
    off = blk[1:-1, [0, 1, 3, 4]].reshape(-1)
    off = off[off != 0]
    assert abs(off.mean()) < 0.01
    assert abs(np.mean(np.abs(off) ** 2) - 1) < 0.01
    assert abs(np.var(off.real) - 0.5) < 0.01 and abs(np.var(off.imag) - 0.5) < 0.01
    # diagonal of diagonal blocks shifted by 4
    dg = blk[:nb - 2, hw][:, np.arange(bs), np.arange(bs)].reshape(-1)
    assert abs(dg.mean() - 4) < 0.02
    t = p.tau
    assert abs(np.mean(np.abs(t) ** 2) - 1) < 0.01


def test_deterministic_and_seed_dependent():
    a = BandedTauPencil.generate(4000, 4, 1, 2, seed=9)
    b = BandedTauPencil.generate(4000, 4, 1, 2, seed=9)
    c = BandedTauPencil.generate(4000, 4, 1, 2, seed=10)
    assert np.array_equal(a.band, b.band) and np.array_equal(a.tau, b.tau)
    assert not np.array_equal(a.band, c.band)
    # the value of an entry depends only on (seed, position): a longer pencil
    # with the same seed shares its leading block rows
    d = BandedTauPencil.generate(8000, 4, 1, 2, seed=9)
    nrow = 3 * 16 * 100
    assert np.array_equal(a.band[:nrow], d.band[:nrow])
