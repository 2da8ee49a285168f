"""CPU emulation of the resident step kernel's data movement (csrc/sgh_resident.hip).

The kernel chains four MFMA contractions per Picard iteration through two LDS operand
buffers, with every extent padded to a multiple of 32. This test replays, in NumPy and
for every lane of every wave, the same tile assignment, A-fragment addresses (cosine
tables in global memory), B-fragment addresses (LDS, [column][k] with row stride 136)
and epilogue stores, and checks for each surface shape that

* every address stays inside its buffer (LDS: 128 × 136 bf16 per buffer; tables:
  pad32(n)² per matrix; surface fields: nx·ny elements);
* every LDS operand entry a stage reads was written by the stage before it, and the
  entries outside the surface hold zero;
* the chained stages compute the 2-D transforms of the solver: T = B_yᵀ X B_x (inverse)
  and S = B_y q B_xᵀ (forward), here in float64 with the kernel's layouts.
"""

import numpy as np
import pytest

WAVE = 64
LD = 136          # kResLd
MAXN = 128        # kResMaxN


def pad32(n):
    return (n + 31) // 32 * 32


class Tiles:
    """ResTiles(RT, CT, wave) of the kernel."""

    def __init__(self, RT, CT, wave, nwaves):
        self.g = max(1, nwaves // RT)
        self.rt = wave // self.g
        self.ct0 = wave % self.g
        self.nct = (CT - self.ct0 + self.g - 1) // self.g if (self.rt < RT and self.ct0 < CT) else 0
        if self.nct == 0:
            self.rt = -1

    def col(self, q, lane):
        return (self.ct0 + self.g * q) * 16 + (lane & 15)

    def row(self, lane, v):
        return self.rt * 16 + 4 * (lane >> 4) + v


def basis(n):
    k = np.arange(n)[:, None]
    i = np.arange(n)[None, :]
    return np.where(k == 0, 1.0, np.sqrt(2.0)) * np.cos(np.pi * k * (2 * i + 1) / (2 * n))


def table(n):
    """The kernel's per-size table: B̂ and B̂ᵀ, zero-padded to pad32(n) (one plane each)."""
    p = pad32(n)
    B = np.zeros((p, p))
    B[:n, :n] = basis(n)
    return B, B.T.copy()


def run_stage(A, src, written, RT, CT, KS, nwaves, epi):
    """One res_stage: every wave, tile and k step; A-fragment rows and k ranges, B reads
    from `src` at [c][k] checked against `written`; epi(r, c, value) per output."""
    lda = A.shape[0]
    for wave in range(nwaves):
        t = Tiles(RT, CT, wave, nwaves)
        if t.rt < 0:
            continue
        for q in range(t.nct):
            acc = np.zeros((WAVE, 4))
            for ks in range(KS):
                for lane in range(WAVE):
                    arow = t.rt * 16 + (lane & 15)
                    k0 = 32 * ks + 8 * (lane >> 4)
                    assert arow < lda and k0 + 8 <= lda
                    c = t.col(q, lane)
                    boff = c * LD + k0
                    assert 0 <= boff and boff + 8 <= MAXN * LD
                    assert written[c, k0:k0 + 8].all(), (c, k0)
                # the 16×16 tile product, lane-wise as the accumulator layout stores it
                for lane in range(WAVE):
                    c = t.col(q, lane)
                    for v in range(4):
                        r = t.row(lane, v)
                        ks_range = slice(32 * ks, 32 * ks + 32)
                        acc[lane, v] += A[r, ks_range] @ src[c, ks_range]
            for lane in range(WAVE):
                for v in range(4):
                    epi(t.row(lane, v), t.col(q, lane), acc[lane, v])


def new_buffer():
    return np.full((MAXN, LD), np.nan), np.zeros((MAXN, LD), dtype=bool)


@pytest.mark.parametrize("nx,ny", [(128, 128), (96, 128), (128, 96), (48, 32), (32, 48),
                                   (64, 96), (2, 1), (17, 33), (100, 7), (128, 2)])
@pytest.mark.parametrize("nwaves", [8])
def test_resident_stage_chain_matches_the_transforms(nx, ny, nwaves):
    rng = np.random.default_rng(nx * 1000 + ny)
    nxp, nyp = pad32(nx), pad32(ny)
    Bx, BxT = table(nx)
    By, ByT = table(ny)

    # ---- forward transform of a physical field q[j][i] (prep writes [j][i], padded) ----
    q = rng.standard_normal((ny, nx))
    buf, wr = new_buffer()
    buf[:nyp, :nxp] = 0.0
    buf[:ny, :nx] = q
    wr[:nyp, :nxp] = True
    # stage 3: Rᵀ = B̂x·qᵀ (rows m: x, cols j: y), tiles tp(nxp/16, nyp/16), stored [m][j]
    nb, nw = new_buffer()

    def store_n(dst, w, r, c, val):
        assert r < MAXN and c < LD - 8
        dst[r, c] = val
        w[r, c] = True
    run_stage(Bx, buf, wr, nxp // 16, nyp // 16, nxp // 32, nwaves,
              lambda r, c, val: store_n(nb, nw, r, c, val))
    # stage 4: S = B̂y·R (rows n, cols m), tiles tm(nyp/16, nxp/16)
    S = np.zeros((nyp, nxp))

    def keep(r, c, val):
        S[r, c] = val
    run_stage(By, nb, nw, nyp // 16, nxp // 16, nyp // 32, nwaves, keep)
    ref = basis(ny) @ q @ basis(nx).T
    np.testing.assert_allclose(S[:ny, :nx], ref, atol=1e-10 * max(1.0, np.abs(ref).max()))
    assert np.abs(S[ny:, :]).max(initial=0) == 0 and np.abs(S[:, nx:]).max(initial=0) == 0

    # ---- inverse transform of a modal field X[n][m], stored [m][n] by the epilogue ----
    X = rng.standard_normal((ny, nx))
    buf, wr = new_buffer()
    Xs = np.zeros((nyp, nxp))
    Xs[:ny, :nx] = X

    def store_t(dst, w, r, c, val):        # res_store_t: element (r, c) at [c][r]
        assert c < MAXN and r < LD - 8
        dst[c, r] = val
        w[c, r] = True
    # the forward epilogue writes every (n, m) of the padded modal tiles
    for wave in range(nwaves):
        t = Tiles(nyp // 16, nxp // 16, wave, nwaves)
        if t.rt < 0:
            continue
        for qq in range(t.nct):
            for lane in range(WAVE):
                for v in range(4):
                    r, c = t.row(lane, v), t.col(qq, lane)
                    store_t(buf, wr, r, c, Xs[r, c])
    # stage 1: P = B̂yᵀ·X (rows j, cols m), tiles tm; stored straight [j][m]
    pb, pw = new_buffer()
    run_stage(ByT, buf, wr, nyp // 16, nxp // 16, nyp // 32, nwaves,
              lambda r, c, val: store_n(pb, pw, r, c, val))
    # stage 2: Tᵀ = B̂xᵀ·Pᵀ (rows i, cols j), tiles tp
    TT = np.zeros((nxp, nyp))

    def keep_t(r, c, val):
        TT[r, c] = val
    run_stage(BxT, pb, pw, nxp // 16, nyp // 16, nxp // 32, nwaves, keep_t)
    ref = basis(ny).T @ X @ basis(nx)
    np.testing.assert_allclose(TT[:nx, :ny].T, ref, atol=1e-10 * max(1.0, np.abs(ref).max()))

    # ---- epilogue offsets stay inside the surface ----
    for wave in range(nwaves):
        for t, rows_are_x in ((Tiles(nxp // 16, nyp // 16, wave, nwaves), True),
                              (Tiles(nyp // 16, nxp // 16, wave, nwaves), False)):
            if t.rt < 0:
                continue
            for qq in range(t.nct):
                for lane in range(WAVE):
                    for v in range(4):
                        r, c = t.row(lane, v), t.col(qq, lane)
                        x, y = (r, c) if rows_are_x else (c, r)
                        if x < nx and y < ny:
                            assert 0 <= y * nx + x < nx * ny


def test_every_padded_tile_has_a_wave():
    # all RT × CT tiles of every stage shape are assigned, each once
    for nwaves in (8, 16):
        for RT in (2, 4, 6, 8):
            for CT in (2, 4, 6, 8):
                seen = []
                for wave in range(nwaves):
                    t = Tiles(RT, CT, wave, nwaves)
                    if t.rt >= 0:
                        seen += [(t.rt, t.ct0 + t.g * q) for q in range(t.nct)]
                        assert t.nct <= (8 if nwaves == 8 else 4)
                assert sorted(seen) == [(r, c) for r in range(RT) for c in range(CT)]
