"""CPU emulation of the resident step kernel's data movement (csrc/sgh_resident.hip).

Every stage of the kernel computes O[a][t] = Σ_k D[a][k]·W[k][t] on the matrix cores: the
field D is read from an LDS operand buffer ([a][k], row stride 136 bf16), the cosine matrix
W from its table in global memory (Wᵀ[t][k], zero-padded to pad32(n)²), and each lane's
four consecutive a of one t are stored as [t][a] into the other buffer. This test replays,
in NumPy and for every lane of every wave, the kernel's tile assignment, its ring of
operand reads (the k steps of all tiles as one sequence, read ahead and clamped to the
wave's last tile), the matrix fragment reads and the stores, and checks for each surface
shape that

* every address stays inside its buffer (LDS: 128 × 136 bf16 per buffer; tables:
  pad32(n)² per matrix);
* every LDS operand entry a stage reads was written by the stage before it, including the
  zero padding;
* the chained stages compute the 2-D transforms of the solver, T = B̂yᵀ X B̂x (inverse) and
  S = B̂y q B̂xᵀ (forward), here in float64 with the kernel's layouts;
* the epilogue offsets y·nx + x stay inside the surface.
"""

import numpy as np
import pytest

WAVE = 64
WAVES = 8         # kResThreads / 64
LD = 136          # kResLd
MAXN = 128        # kResMaxN


def pad32(n):
    return (n + 31) // 32 * 32


def ring(KS):
    """Ring size and read-ahead distance of res_tiles<KS>."""
    R = {1: 2, 2: 4, 3: 3, 4: 4}[KS]
    assert (2 * KS) % R == 0
    return R, R - 1


class Tiles:
    """ResTiles(NT, NA, wave) of the kernel."""

    def __init__(self, NT, NA, wave):
        self.g = max(1, WAVES // NT)
        self.tt = wave // self.g
        self.a0 = wave % self.g
        self.nq = (NA - self.a0 + self.g - 1) // self.g if (self.tt < NT and self.a0 < NA) else 0
        if self.nq == 0:
            self.tt = -1

    def arow(self, q, lane):
        return (self.a0 + self.g * q) * 16 + (lane & 15)

    def aout(self, q, lane):
        return (self.a0 + self.g * q) * 16 + 4 * (lane >> 4)

    def tcol(self, lane):
        return self.tt * 16 + (lane & 15)


def basis(n):
    k = np.arange(n)[:, None]
    i = np.arange(n)[None, :]
    return np.where(k == 0, 1.0, np.sqrt(2.0)) * np.cos(np.pi * k * (2 * i + 1) / (2 * n))


def padded(M):
    p = pad32(M.shape[0])
    out = np.zeros((p, p))
    out[:M.shape[0], :M.shape[1]] = M
    return out


def run_stage(WT, src, written, NT, NA, KS, dst=None, dst_w=None):
    """One res_stage. WT: the table Wᵀ[t][k] (pad32 square). Returns O[a][t] for the
    padded tiles and, with dst, stores it as [t][a] the way res_store does."""
    R, AH = ring(KS)
    p = WT.shape[0]
    out = {}
    for wave in range(WAVES):
        t = Tiles(NT, NA, wave)
        if t.tt < 0:
            continue
        # matrix fragments, k steps beyond KS clamped to the last one
        for ks in range(4):
            kk = min(ks, KS - 1)
            for lane in range(WAVE):
                row, k0 = t.tcol(lane), 32 * kk + 8 * (lane >> 4)
                assert row < p and k0 + 8 <= p
        # operand reads: the sequence s = q·KS + ks for q < nq, then AH steps beyond
        for s in range(t.nq * KS + AH):
            q, ks = divmod(s, KS)
            qq = min(q, t.nq - 1)
            for lane in range(WAVE):
                a, k0 = t.arow(qq, lane), 32 * ks + 8 * (lane >> 4)
                off = a * LD + k0
                assert 0 <= off and off + 8 <= MAXN * LD and k0 + 8 <= LD
                assert written[a, k0:k0 + 8].all(), (a, k0)
        for q in range(t.nq):
            for lane in range(WAVE):
                tc = t.tcol(lane)
                for v in range(4):
                    a = t.aout(q, lane) + v
                    val = src[a, :32 * KS] @ WT[tc, :32 * KS]
                    out[a, tc] = val
                    if dst is not None:
                        assert tc < MAXN and a < LD - 8
                        dst[tc, a] = val
                        dst_w[tc, a] = True
    return out


def new_buffer():
    return np.full((MAXN, LD), np.nan), np.zeros((MAXN, LD), dtype=bool)


@pytest.mark.parametrize("nx,ny", [(128, 128), (96, 128), (128, 96), (48, 32), (32, 48),
                                   (64, 96), (2, 1), (17, 33), (100, 7), (128, 2)])
def test_resident_stage_chain_matches_the_transforms(nx, ny):
    rng = np.random.default_rng(nx * 1000 + ny)
    nxp, nyp = pad32(nx), pad32(ny)
    Bx, By = padded(basis(nx)), padded(basis(ny))
    KX, KY = nxp // 32, nyp // 32

    # ---- forward: q stored [i][j] (prep writes the padded surface) ----
    q = rng.standard_normal((ny, nx))
    buf, wr = new_buffer()
    buf[:nxp, :nyp] = 0.0
    buf[:nx, :ny] = q.T
    wr[:nxp, :nyp] = True
    # fwd1: R[i][n] = Σ_j q[i][j] B̂y[n][j]; Wᵀ[n][j] = B̂y; stored [n][i]
    b1, w1 = new_buffer()
    run_stage(By, buf, wr, nyp // 16, nxp // 16, KY, b1, w1)
    # fwd2: S[n][m] = Σ_i R[n][i] B̂x[m][i]; Wᵀ[m][i] = B̂x
    out = run_stage(Bx, b1, w1, nxp // 16, nyp // 16, KX)
    S = np.zeros((nyp, nxp))
    for (a, t), v in out.items():
        S[a, t] = v
    ref = basis(ny) @ q @ basis(nx).T                       # [n][m]
    np.testing.assert_allclose(S[:ny, :nx], ref, atol=1e-10 * max(1.0, np.abs(ref).max()))
    assert np.abs(S[ny:, :]).max(initial=0) == 0 and np.abs(S[:, nx:]).max(initial=0) == 0
    assert len(out) == nxp * nyp                             # every padded output, once

    # ---- inverse: X as the forward epilogue stores it ([m][n], zero padding) ----
    X = rng.standard_normal((ny, nx))
    buf, wr = new_buffer()
    buf[:nxp, :nyp] = 0.0
    buf[:nx, :ny] = X.T
    wr[:nxp, :nyp] = True
    # inv1: V[m][j] = Σ_n X[m][n] B̂y[n][j]; Wᵀ[j][n] = B̂yᵀ; stored [j][m]
    b1, w1 = new_buffer()
    run_stage(By.T.copy(), buf, wr, nyp // 16, nxp // 16, KY, b1, w1)
    # inv2: T[j][i] = Σ_m V[j][m] B̂x[m][i]; Wᵀ[i][m] = B̂xᵀ; stored [i][j]
    b0, w0 = new_buffer()
    run_stage(Bx.T.copy(), b1, w1, nxp // 16, nyp // 16, KX, b0, w0)
    ref = basis(ny).T @ X @ basis(nx)                        # [j][i]
    np.testing.assert_allclose(b0[:nx, :ny].T, ref, atol=1e-10 * max(1.0, np.abs(ref).max()))
    assert w0[:nxp, :nyp].all()      # the next forward stage reads all of it

    # ---- epilogue offsets (second-stage tiles: t along x, a along y) ----
    for wave in range(WAVES):
        t = Tiles(nxp // 16, nyp // 16, wave)
        if t.tt < 0:
            continue
        for qq in range(t.nq):
            for lane in range(WAVE):
                x = t.tcol(lane)
                for v in range(4):
                    y = t.aout(qq, lane) + v
                    if x < nx and y < ny:
                        assert 0 <= y * nx + x < nx * ny


def test_every_padded_tile_has_a_wave():
    # all NT × NA tiles of every stage shape are assigned, each once
    for NT in (1, 2, 4, 6, 8):
        for NA in (1, 2, 4, 6, 8):
            seen = []
            for wave in range(WAVES):
                t = Tiles(NT, NA, wave)
                if t.tt >= 0:
                    seen += [(t.tt, t.a0 + t.g * q) for q in range(t.nq)]
                    assert t.nq <= 8
            assert sorted(seen) == [(r, c) for r in range(NT) for c in range(NA)]


# ---- teams: a cuboid's step on TP workgroups (members) ------------------------------
KRES_BUF = MAXN * LD                          # bf16 of one hi or lo operand buffer
XCH_FIELD = 2 * KRES_BUF                      # bf16 of a hi+lo exchange field
XCH_NORM = 2 * XCH_FIELD * 2                  # byte offset of the per-tile norm sums
TEAM_MAX, TILES = 8, 64
NORM_SLOTS = 4 * TILES                        # one per 4-row group of a tile
XCH_BYTES = (XCH_NORM + 4 * 3 * NORM_SLOTS + 8 * TEAM_MAX + 8 * (MAXN // 16) + 255) // 256 * 256


class TeamTiles(Tiles):
    """ResTiles(NT, NA, wave, toff, aoff): t tiles numbered from toff, a tiles from aoff."""

    def __init__(self, NT, NA, wave, toff, aoff):
        super().__init__(NT, NA, wave)
        if self.tt >= 0:
            self.tt += toff
        self.a0 += aoff


def run_member_stage(WT, src, written, NT, NA, KS, toff, aoff, dst=None, dst_w=None):
    """run_stage for one member's tiles (the kernel's ResTiles with offsets)."""
    R, AH = ring(KS)
    p = WT.shape[0]
    out = {}
    for wave in range(WAVES):
        t = TeamTiles(NT, NA, wave, toff, aoff)
        if t.tt < 0:
            continue
        for lane in range(WAVE):
            assert t.tcol(lane) < p
        for s in range(t.nq * KS + AH):
            q, ks = divmod(s, KS)
            qq = min(q, t.nq - 1)
            for lane in range(WAVE):
                a, k0 = t.arow(qq, lane), 32 * ks + 8 * (lane >> 4)
                assert 0 <= a * LD + k0 and a * LD + k0 + 8 <= MAXN * LD
                assert written[a, k0:k0 + 8].all(), (a, k0)
        for q in range(t.nq):
            for lane in range(WAVE):
                tc = t.tcol(lane)
                for v in range(4):
                    a = t.aout(q, lane) + v
                    val = src[a, :32 * KS] @ WT[tc, :32 * KS]
                    out[a, tc] = val
                    if dst is not None:
                        dst[tc, a] = val
                        dst_w[tc, a] = True
    return out


def exchange(members, field, nxp):
    """The kernel's exchange: every member's column band of its buffer (rows < nxp) to the
    exchange field by 16-byte units, then rows < nxp of the field into every member's
    buffer by LDS DMA. members: list of (a_lo, a_hi, buf, written)."""
    g = np.full(KRES_BUF, np.nan)
    gw = np.zeros(KRES_BUF, dtype=bool)
    for a_lo, a_hi, buf, wr in members:
        ur = 2 * (a_hi - a_lo)
        per = nxp * ur
        for u in range(per):          # the hi field; lo has the same offsets + KRES_BUF
            r = u // ur
            off = r * LD + 16 * a_lo + 8 * (u - r * ur)
            assert off + 8 <= KRES_BUF
            assert 2 * (field + KRES_BUF + off + 8) <= XCH_NORM
            flat, fw = buf.reshape(-1), wr.reshape(-1)
            assert fw[off:off + 8].all(), (a_lo, r, off)
            assert not gw[off:off + 8].any(), "two members store the same unit"
            g[off:off + 8] = flat[off:off + 8]
            gw[off:off + 8] = True
    units = nxp * (LD // 8)
    assert 8 * units <= KRES_BUF
    out = []
    for _ in members:
        b, w = new_buffer()
        b.reshape(-1)[:8 * units] = g[:8 * units]
        w.reshape(-1)[:8 * units] = gw[:8 * units]
        out.append((b, w))
    return out


def assert_epilogue_writes_exactly_the_band(lo, hi, w, nxp):
    """Inside the Picard loop the epilogues store their tiles straight to the exchange
    field (same layout) instead of the exchange copying the band out of LDS: that stores
    the same units only if the epilogue writes every unit of the member's band (rows < nxp,
    columns [16·lo, 16·hi)) and nothing outside it (another member's band)."""
    band = np.zeros_like(w)
    band[:nxp, 16 * lo:16 * hi] = True
    assert (w == band).all(), (lo, hi)
    assert 2 * (XCH_FIELD + KRES_BUF + (nxp - 1) * LD + 16 * hi) <= XCH_NORM


@pytest.mark.parametrize("nx,ny", [(128, 128), (96, 128), (128, 96), (48, 64), (32, 33),
                                   (100, 70), (17, 128)])
def test_resident_team_stages_match_the_transforms(nx, ny):
    rng = np.random.default_rng(nx * 7 + ny)
    nxp, nyp = pad32(nx), pad32(ny)
    NTB, NAY = nxp // 16, nyp // 16
    KX, KY = nxp // 32, nyp // 32
    Bx, By = padded(basis(nx)), padded(basis(ny))
    for TP in range(1, min(TEAM_MAX, NAY) + 1):
        bands = [(m * NAY // TP, (m + 1) * NAY // TP) for m in range(TP)]
        assert all(hi > lo for lo, hi in bands) and bands[-1][1] == NAY
        # prep rows e in [16·a_lo·nxp, 16·a_hi·nxp) partition the padded surface
        rows = sorted(r for lo, hi in bands for r in range(16 * lo, 16 * hi))
        assert rows == list(range(nyp))

        # ---- forward: every member holds the whole q (after the prep's exchange) ----
        q = rng.standard_normal((ny, nx))
        S = np.zeros((nyp, nxp))
        X_members = []
        for (lo, hi) in bands:
            buf, wr = new_buffer()
            buf[:nxp, :nyp] = 0.0
            buf[:nx, :ny] = q.T
            wr[:nxp, :nyp] = True
            b1, w1 = new_buffer()
            run_member_stage(By, buf, wr, hi - lo, NTB, KY, lo, 0, b1, w1)
            out = run_member_stage(Bx, b1, w1, NTB, hi - lo, KX, 0, lo)
            for (a, t), v in out.items():
                assert lo * 16 <= a < hi * 16
                S[a, t] = v
            # the epilogue's X into the member's buffer 0 as [m][n], its own n band only
            b0, w0 = new_buffer()
            for (a, t), v in out.items():
                b0[t, a] = v
                w0[t, a] = True
            assert_epilogue_writes_exactly_the_band(lo, hi, w0, nxp)
            X_members.append((lo, hi, b0, w0))
        ref = basis(ny) @ q @ basis(nx).T
        np.testing.assert_allclose(S[:ny, :nx], ref, atol=1e-10 * max(1.0, np.abs(ref).max()))
        full = exchange(X_members, 0, nxp)
        for b, w in full:
            assert w[:nxp, :nyp].all()       # the inverse's first stage reads all of it
            np.testing.assert_array_equal(b[:nxp, :nyp], S[:nyp, :nxp].T)

        # ---- inverse from the exchanged X; evaporation epilogue's q band; exchange ----
        X = rng.standard_normal((ny, nx))
        T = np.zeros((nyp, nxp))
        q_members = []
        for (lo, hi) in bands:
            buf, wr = new_buffer()
            buf[:nxp, :nyp] = 0.0
            buf[:nx, :ny] = X.T
            wr[:nxp, :nyp] = True
            b1, w1 = new_buffer()
            run_member_stage(By.T.copy(), buf, wr, hi - lo, NTB, KY, lo, 0, b1, w1)
            b0, w0 = new_buffer()
            out = run_member_stage(Bx.T.copy(), b1, w1, NTB, hi - lo, KX, 0, lo, b0, w0)
            for (a, t), v in out.items():
                T[a, t] = v
            assert_epilogue_writes_exactly_the_band(lo, hi, w0, nxp)
            q_members.append((lo, hi, b0, w0))
        ref = basis(ny).T @ X @ basis(nx)
        np.testing.assert_allclose(T[:ny, :nx], ref, atol=1e-10 * max(1.0, np.abs(ref).max()))
        for b, w in exchange(q_members, XCH_FIELD, nxp):
            assert w[:nxp, :nyp].all()
            np.testing.assert_array_equal(b[:nxp, :nyp], T[:nyp, :nxp].T)

        # ---- per-tile norm slots and work-slot indices: disjoint, in range ----
        slots, units = [], []
        for (lo, hi) in bands:
            for wave in range(WAVES):
                t = TeamTiles(NTB, hi - lo, wave, 0, lo)
                if t.tt < 0:
                    continue
                for qq in range(t.nq):
                    a = t.a0 + t.g * qq
                    assert lo <= a < hi
                    # the row group of lane l is l >> 4: lanes 15, 31, 47, 63 store
                    slots += [(a * 8 + t.tt) * 4 + (lane >> 4) for lane in (15, 31, 47, 63)]
                    units += [(a * NTB + t.tt) * 64 + lane for lane in range(WAVE)]
        assert sorted(slots) == sorted((a * 8 + tt) * 4 + g for a in range(NAY)
                                       for tt in range(NTB) for g in range(4))
        assert all(0 <= s < NORM_SLOTS for s in slots)
        # the sums read lane a·8 + t's four slots (16 bytes at 16·lane) for a < NAY, t < NTB
        read = sorted(4 * (a * 8 + tt) + g for a in range(NAY) for tt in range(NTB)
                      for g in range(4))
        assert read == sorted(slots)
        assert len(set(units)) == len(units) and max(units) < MAXN * MAXN // 4
        # T_max per member, then the laser power per 16-row band (member 0 sums them in
        # band order)
        assert XCH_NORM + 4 * 3 * NORM_SLOTS + 8 * TEAM_MAX + 8 * NAY <= XCH_BYTES
        assert NAY <= MAXN // 16


def test_team_work_list_entries_round_trip():
    # res_entry(c, tp, p) = c·64 + (tp − 1)·8 + p, decoded as c = e >> 6, tp = ((e >> 3) & 7) + 1,
    # p = e & 7
    for c in (0, 1, 63, 1000):
        for tp in range(1, TEAM_MAX + 1):
            for p in range(tp):
                e = c * 64 + (tp - 1) * 8 + p
                assert (e >> 6, ((e >> 3) & 7) + 1, e & 7) == (c, tp, p)
