"""ORACLE -- test infrastructure only. Never imported by the product package.

FP64 NumPy/SciPy restatement of the reference's spectral Galerkin ETD1 + Picard step
(``SpectralSolver.step``, ``src/spectral_galerkin_heat/solvers/spectral.py:219-543``),
in grid mode and in box mode (laser-following fine sub-box), with every physics kernel
it calls. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may use it, as the
checker the HIP path is compared against.

Function -> reference source it restates, and what pins it (tests/test_oracle_*.py):

=============================  ==================================================  =========
function                       reference                                           pinned by
=============================  ==================================================  =========
c_coef                         physics/spectral_helpers.py:26-46                   closed form
dct2 / idct2                   physics/spectral_cpu_kernels.py:374-389             brute-force cosine sums, DCT(1)=n^{3/2}, round trip
precompute_K_KK                physics/spectral_state.py:374-421                   K[0]=1, KK[0]=dt/(ρc), exp closed form, zyx order
OracleGrid                     physics/spectral_state.py:71-143                    cell centres, IC → T0 surface, modal sum
update_modes_etd1 & co.        physics/spectral_cpu_kernels.py:83-113              hand-computed values
evaporation_flux               physics/spectral_cpu_kernels.py:144-162             re-derived Arrhenius law, gate, monotone
latent_source                  physics/spectral_cpu_kernels.py:115-140             band, sign, magnitude, clamp
latent_source_model            core/properties.py:375-430                          equals latent_source for constant ρ
shift_flux                     physics/spectral_cpu_kernels.py:281-396             half-pixel ramp; scipy.ndimage.shift(order=1)
super_gaussian_flux            core/laser.py:42-53                                 peak = I₀, integral I₀πr²/2
gaussian_cell_integrated_flux  core/laser.py:65-90                                 power = A·P for sub-grid beams
correction_source              physics/spectral_cpu_kernels.py:165-252             np.gradient divergence form
assemble_property_correction   physics/spectral_ops.py:105-170                     null-exactness, steady reduction
OracleSolver.step              solvers/spectral.py:219-543                         laser-off → T0, P_laser ≈ AP, energy balance,
                                                                                   linear limit vs Eagar–Tsai closed form
OracleLinearSolver.step        solvers/spectral_cpu_linear.py:98-162               Eagar–Tsai closed form, superposition
reconstruct_temperature_DCT    physics/spectral_helpers.py:147-207                 direct modal sum at nodes
reconstruct_at_points          physics/spectral_helpers.py:209-255                 = cell-centred volume at centres, = DCT-I at nodes
subgrid_indices                physics/spectral_helpers.py:49-83                   reference test vectors (test_spectral_helpers.py:32-54)
OracleFineMesh                 physics/spectral_state.py:166-239                   brute-force cosines at fine centres
reconstruct_box / project_box  physics/spectral_ops.py:53-74,173-179               project∘reconstruct = I (full box, r = 1..3);
                                                                                   r = 1 box = cell-centred volume transforms
shift_3d_xy                    physics/spectral_cpu_kernels.py:302-322             scipy.ndimage.shift nearest/constant;
                                                                                   lab-frame agreement with the window slide
box mode of OracleSolver.step  solvers/spectral.py:286-534, spectral_ops.py:182-306  r = 1 full box reproduces grid mode
                                                                                   (test_latent_grid_mode.py:58-97); latent sign
box_face_peaks                 solvers/spectral.py:545-586                         warning raised when the pool reaches the floor
compute_L2_structured          io_utils/compute_L2_error.py:436-531                closed form
=============================  ==================================================  =========

The multi-step nonlinear trajectory (warm-start order, lagged bottom convection, the
grid-mode latent replay) has no golden vector in the reference; it is pinned by the
invariants above plus the exact energy balance of the mean mode, and is otherwise
"parity unpinned" against the reference itself.
"""

from __future__ import annotations

import ast
import math

import numpy as np
import scipy.fft
import scipy.special

# ----------------------------------------------------------------------------------
# Basis, transforms, propagators
# ----------------------------------------------------------------------------------


def c_coef(n, L):
    """Cosine-basis norms: sqrt(2/L), and sqrt(1/L) for mode 0 (spectral_helpers.py:26-46)."""
    C = np.sqrt(2.0 / L) * np.ones(n)
    C[0] = np.sqrt(1.0 / L)
    return C


# scipy.fft worker threads of the oracle's transforms: -1 = every core (the reference's
# setting, spectral_cpu_kernels.py:374-389); the timing of a one-core baseline sets 1.
FFT_WORKERS = -1


def dct2(q):
    """Ortho DCT-II over every axis (spectral_cpu_kernels.py:374-381)."""
    arr = np.ascontiguousarray(q)
    return scipy.fft.dctn(arr, type=2, norm="ortho", axes=tuple(range(arr.ndim)),
                          workers=FFT_WORKERS)


def idct2(a):
    """Ortho DCT-III over every axis, the inverse of dct2 (spectral_cpu_kernels.py:383-389)."""
    arr = np.ascontiguousarray(a)
    return scipy.fft.dctn(arr, type=3, norm="ortho", axes=tuple(range(arr.ndim)),
                          workers=FFT_WORKERS)


def precompute_K_KK(k, rho, Cp, dt, n_zyx, L_zyx, dtype=np.float64):
    """E = exp(-αλΔt), KK = φ₁/(ρc) with KK[0,0,0] = Δt/(ρc) (spectral_state.py:374-421)."""
    nz, ny, nx = n_zyx
    Lz, Ly, Lx = L_zyx
    kz, ky, kx = [np.pi * np.arange(n) / L for n, L in zip((nz, ny, nx), (Lz, Ly, Lx))]
    alpha = k / (rho * Cp)
    rho_Cp = rho * Cp
    denom = alpha * (kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2)
    K = np.exp(-denom * dt).astype(dtype)
    denom[0, 0, 0] = 1.0
    phi_1 = (K - 1.0) / (-denom)
    phi_1[0, 0, 0] = dt
    KK = (phi_1 / rho_Cp).astype(dtype)
    return K, KK


class OracleGrid:
    """Cell-centred grid constants (spectral_state.py:98-143)."""

    def __init__(self, size, n):
        (Lx, Ly, Lz), (nx, ny, nz) = size, n
        self.Lx, self.Ly, self.Lz = float(Lx), float(Ly), float(Lz)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx, self.dy, self.dz = self.Lx / nx, self.Ly / ny, self.Lz / nz
        self.x = (np.arange(nx) + 0.5) * self.dx
        self.y = (np.arange(ny) + 0.5) * self.dy
        self.z = (np.arange(nz) + 0.5) * self.dz
        self.C = (c_coef(nx, Lx), c_coef(ny, Ly), c_coef(nz, Lz))
        self.dct_scale = (self.dx * self.dy) * np.sqrt((nx * ny) / (Lx * Ly))
        self.recon_scale = np.sqrt(nx * ny) / np.sqrt(Lx * Ly)
        sign = np.power(-1.0, np.arange(nz))
        self.Cp_top = (self.C[2] * sign)[:, None, None]
        self.Cp_bot = self.C[2][:, None, None].copy()
        self.sqrt_dV = np.sqrt(self.dx * self.dy * self.dz)


def reconstruct_surface(a, g):
    """Top surface (z = Lz): Σ_p C_p(−1)^p a_p then 2-D IDCT (spectral_ops.py:210-219)."""
    A = np.einsum("p,pij->ij", g.Cp_top[:, 0, 0], a, optimize=True)
    return g.recon_scale * idct2(A)


def reconstruct_bottom(a, g):
    """Bottom surface (z = 0): Σ_p C_p a_p then 2-D IDCT (spectral_ops.py:222-231)."""
    A = np.einsum("p,pij->ij", g.Cp_bot[:, 0, 0], a, optimize=True)
    return g.recon_scale * idct2(A)


def reconstruct_volume(a, g):
    """T = IDCT(a)/sqrt(dV) on the cell-centred grid (spectral_ops.py:83-91)."""
    return idct2(a) / g.sqrt_dV


def project_volume(f, g):
    """modes = sqrt(dV)·DCT(f) (spectral_ops.py:94-102)."""
    return dct2(f) * g.sqrt_dV


# ----------------------------------------------------------------------------------
# Box mode: the laser-following fine sub-box
# ----------------------------------------------------------------------------------


def subgrid_indices(pos, d, n_total, n_box):
    """(start, end, offset) of an n_box window centred on ``pos`` (spectral_helpers.py:49-83).

    The centre cell is Python's round() (half to even) of pos/d clamped to the grid; the
    window start is clamped to [0, max(0, n_total - n_box)].
    """
    centre = round(max(0.0, min(pos / d, n_total - 1)))
    start = max(0, min(centre - n_box // 2, max(0, n_total - n_box)))
    return start, start + n_box, max(0, min(centre - start, n_box - 1))


class OracleFineMesh:
    """Fine sub-box bases and history (FineMeshState, spectral_state.py:166-239).

    Fine spacing d/r per axis; x and y cover the whole domain with a sliding window of
    n_box cells, z covers the top ``Lz_box`` of the domain (cell centres shifted by
    Lz − Lz_box) and does not slide. B_full[a][k, i] = C_k·cos(π·k·coord_i/L_a).
    """

    def __init__(self, size, n, refinement, box_size):
        Lx, Ly, Lz = (float(v) for v in size)
        nx, ny, nz = (int(v) for v in n)
        r = int(refinement)
        self.refinement = r
        d = (Lx / nx, Ly / ny, Lz / nz)
        self.d_fine = [dd / r for dd in d]
        Lbx, Lby, Lbz = (float(v) for v in box_size)
        self.n_totals = [int(np.ceil(L / df)) for L, df in zip((Lx, Ly, Lbz), self.d_fine)]
        coords = [(np.arange(m) + 0.5) * df for m, df in zip(self.n_totals, self.d_fine)]
        coords[2] = (Lz - Lbz) + coords[2]
        modes, lengths = (nx, ny, nz), (Lx, Ly, Lz)
        self.B_full = [c_coef(modes[a], lengths[a])[:, None]
                       * np.cos(np.pi * np.arange(modes[a])[:, None] * coords[a][None, :]
                                / lengths[a]) for a in range(3)]
        self.n_box = [min(int(np.ceil(Lb / df)), m)
                      for Lb, df, m in zip((Lbx, Lby, Lbz), self.d_fine, self.n_totals)]
        self.start = [0, 0, 0]
        self.dV_fine = self.d_fine[0] * self.d_fine[1] * self.d_fine[2]
        self.T_prev = None
        self.Q_prev = None

    def update(self, x, y):
        """Slide the x and y windows to the laser position (spectral_state.py:227-235)."""
        for a, pos in ((0, x), (1, y)):
            self.start[a] = subgrid_indices(float(pos), self.d_fine[a], self.n_totals[a],
                                            self.n_box[a])[0]

    def bases(self):
        """The active window bases (x, y, z), each (modes, n_box)."""
        return [self.B_full[a][:, self.start[a]:self.start[a] + self.n_box[a]] for a in range(3)]


def reconstruct_box(a, fm):
    """T on the fine box cells: Σ a·B_z B_y B_x (spectral_ops.py:173-179)."""
    Bx, By, Bz = fm.bases()
    return np.einsum("ZYX,Zz,Yy,Xx->zyx", a, Bz, By, Bx, optimize=True)


def project_box(f, fm):
    """Modes of a box field: dV_fine·Σ f·B_z B_y B_x (spectral_ops.py:53-74)."""
    Bx, By, Bz = fm.bases()
    return np.einsum("zyx,Zz,Yy,Xx->ZYX", f, Bz, By, Bx, optimize=True) * fm.dV_fine


def shift_3d_xy(field, sy, sx, nearest):
    """Order-1 (y, x) shift of every z slice (_shift_xy_3d, spectral_cpu_kernels.py:302-322)."""
    return np.stack([shift_2d(field[z], sy, sx, nearest=nearest) for z in range(field.shape[0])])


def box_face_peaks(Q, fm):
    """Faces the latent source must not reach, as in ``_check_latent_box_truncation``
    (spectral.py:545-586): returns the names whose max |Q| exceeds 1e-2 of the box peak.
    The floor is always checked; x/y window edges only when the window is narrower than
    the fine grid."""
    peak = float(np.max(np.abs(Q)))
    if peak <= 0.0:
        return []
    faces = {"floor (depth)": Q[0, :, :]}
    if fm.n_box[0] < fm.n_totals[0]:
        faces["x window edge"] = Q[:, :, (0, -1)]
    if fm.n_box[1] < fm.n_totals[1]:
        faces["y window edge"] = Q[:, (0, -1), :]
    return [k for k, f in faces.items() if float(np.max(np.abs(f))) > 1e-2 * peak]


# ----------------------------------------------------------------------------------
# Pointwise kernels
# ----------------------------------------------------------------------------------


def update_modes_etd1(aK, KK, Cp, B, out):
    """out = aK + KK·Cp[p]·B (spectral_cpu_kernels.py:83-91)."""
    out[...] = aK + KK * Cp * B[None, :, :]


def add_source_term_modes(a, KK, Q):
    """a += KK·Q (spectral_cpu_kernels.py:93-103)."""
    a += KK * Q


def add_bottom_surface_source(a, KK, Cb, B):
    """a += KK·Cb[p]·B (spectral_cpu_kernels.py:105-113)."""
    a += KK * Cb * B[None, :, :]


def evaporation_flux(T, P0, T_boil, dH, R_v, T_liq):
    """0.82·ΔH·P₀/√(2πR_v) · T^-½ · exp(ΔH/(R_v T_b)·(1 − T_b/T)) above T_liq, else 0
    (spectral_cpu_kernels.py:144-162)."""
    f1 = 0.82 * dH * P0 / np.sqrt(2 * np.pi * R_v)
    f2 = dH / (R_v * T_boil)
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        q = f1 * (1.0 / np.sqrt(T)) * np.exp(f2 * (1.0 - T_boil / T))
    return np.where(T < T_liq, 0.0, q)


def _fl(T, T_S, inv_band):
    return np.clip((T - T_S) * inv_band, 0.0, 1.0)


def latent_source(T, Tp, T_S, T_L, rho, L, dt):
    """Q = −ρL(f_l(T) − f_l(T_prev))/dt with clamped f_l (spectral_cpu_kernels.py:115-140)."""
    inv_band = 1.0 / (T_L - T_S)
    return (-rho * L / dt) * (_fl(T, T_S, inv_band) - _fl(Tp, T_S, inv_band))


def super_gaussian_area_factor(order):
    """f(n) = π·2^(1−2/n)·Γ(2/n)/n (laser.py:34-40)."""
    n = float(order)
    return math.pi * 2.0 ** (1.0 - 2.0 / n) * math.gamma(2.0 / n) / n


def super_gaussian_flux(x, y, x0, y0, rx, ry, order, I0):
    """I₀·exp(−2ρⁿ) at cell centres (laser.py:42-53)."""
    ddx = (x[None, :] - x0) / rx
    ddy = (y[:, None] - y0) / ry
    rho_sq = ddx * ddx + ddy * ddy
    n = float(order)
    shape = np.exp(-2.0 * rho_sq) if n == 2.0 else np.exp(-2.0 * rho_sq ** (n / 2.0))
    return I0 * shape


def gaussian_cell_integrated_flux(x, y, x0, y0, rx, ry, hx, hy, I0):
    """Cell-mean Gaussian via 1-D erf factors (laser.py:65-90)."""
    s2 = math.sqrt(2.0)
    nx_ = rx * math.sqrt(math.pi / 8.0)
    ny_ = ry * math.sqrt(math.pi / 8.0)
    ax = (x - x0 - hx / 2.0) * s2 / rx
    bx = (x - x0 + hx / 2.0) * s2 / rx
    ay = (y - y0 - hy / 2.0) * s2 / ry
    by = (y - y0 + hy / 2.0) * s2 / ry
    mx = nx_ * (scipy.special.erf(bx) - scipy.special.erf(ax)) / hx
    my = ny_ * (scipy.special.erf(by) - scipy.special.erf(ay)) / hy
    return I0 * (my[:, None] * mx[None, :])


def _shift_axis_tables(n, shift):
    """Clamped indices, bilinear weights, in-range mask (spectral_cpu_kernels.py:325-348)."""
    o = np.arange(n, dtype=np.float64)
    p = o - float(shift)
    i0 = np.floor(p).astype(np.int64)
    t = p - i0
    inr = (p >= 0.0) & (p <= n - 1)
    return np.clip(i0, 0, n - 1), np.clip(i0 + 1, 0, n - 1), 1.0 - t, t, inr


def shift_2d(field, sy, sx, nearest=False, cval=0.0):
    """Order-1 ndimage shift of a 2-D field, ``out[o] = in(o − s)`` (:299-322, :351-371)."""
    i0y, i1y, w0y, w1y, iny = _shift_axis_tables(field.shape[0], sy)
    i0x, i1x, w0x, w1x, inx = _shift_axis_tables(field.shape[1], sx)
    f = field
    val = (w0y[:, None] * (w0x[None, :] * f[i0y][:, i0x] + w1x[None, :] * f[i0y][:, i1x])
           + w1y[:, None] * (w0x[None, :] * f[i1y][:, i0x] + w1x[None, :] * f[i1y][:, i1x]))
    if not nearest:
        val = np.where(iny[:, None] & inx[None, :], val, cval)
    return val


def shift_flux(field, shift, dx, dy):
    """Translate a surface flux by ``shift = (Δx, Δy)`` metres (spectral_cpu_kernels.py:392-396)."""
    sx_m, sy_m = shift
    return shift_2d(field, sy_m / dy, sx_m / dx)


def _grad(f, d, axis):
    """numpy.gradient stencil, edge_order 1: central inside, one-sided at the faces."""
    return np.gradient(f, d, axis=axis)


def correction_source(T, kp, ap, Tprev, dt, dx, dy, dz):
    """f = ∇·(k'∇T) − a'(T − T_prev)/dt (spectral_cpu_kernels.py:165-252)."""
    gx = kp * _grad(T, dx, 2)
    gy = kp * _grad(T, dy, 1)
    gz = kp * _grad(T, dz, 0)
    div = _grad(gx, dx, 2) + _grad(gy, dy, 1) + _grad(gz, dz, 0)
    return div - ap * ((T - Tprev) * (1.0 / dt))


# ----------------------------------------------------------------------------------
# Temperature-dependent properties (core/properties.py, core/property_expr.py)
# ----------------------------------------------------------------------------------

_FUNCS = {
    "sqrt": np.sqrt, "exp": np.exp, "log": np.log, "log2": np.log2, "log10": np.log10,
    "sin": np.sin, "cos": np.cos, "tan": np.tan, "asin": np.arcsin, "acos": np.arccos,
    "atan": np.arctan, "atan2": np.arctan2, "sinh": np.sinh, "cosh": np.cosh,
    "tanh": np.tanh, "asinh": np.arcsinh, "acosh": np.arccosh, "atanh": np.arctanh,
    "fabs": np.abs, "ceil": np.ceil, "floor": np.floor, "trunc": np.trunc,
    "log1p": np.log1p, "expm1": np.expm1,
}


def eval_expr(src, T):
    """Evaluate a property expression string at T (property_expr.py:124-147, 279-283)."""
    s = str(src).strip()
    try:
        return T * 0.0 + float(s)
    except ValueError:
        pass

    def ev(node):
        if isinstance(node, ast.Constant):
            return node.value
        if isinstance(node, ast.Name):
            return T if node.id == "T" else {"pi": math.pi, "e": math.e}[node.id]
        if isinstance(node, ast.BinOp):
            a, b = ev(node.left), ev(node.right)
            return {ast.Add: lambda: a + b, ast.Sub: lambda: a - b, ast.Mult: lambda: a * b,
                    ast.Div: lambda: a / b, ast.Pow: lambda: a ** b}[type(node.op)]()
        if isinstance(node, ast.UnaryOp):
            v = ev(node.operand)
            return -v if isinstance(node.op, ast.USub) else v
        if isinstance(node, ast.Call):
            return _FUNCS[node.func.id](*(ev(a) for a in node.args))
        raise ValueError(f"unsupported node {type(node).__name__}")

    return ev(ast.parse(s, mode="eval").body)


class OracleProps:
    """Blended k(T), ρ(T), c(T) from six branch strings (properties.py:79-108, 318-395)."""

    def __init__(self, k_s, k_l, rho_s, rho_l, c_s, c_l, T_sol, T_liq):
        self.src = dict(k_s=k_s, k_l=k_l, rho_s=rho_s, rho_l=rho_l, c_s=c_s, c_l=c_l)
        self.T_sol, self.T_liq = float(T_sol), float(T_liq)

    @classmethod
    def from_model(cls, model):
        return cls(model.k.solid.source, model.k.liquid.source, model.rho.solid.source,
                   model.rho.liquid.source, model.c.solid.source, model.c.liquid.source,
                   model.k.T_solidus, model.k.T_liquidus)

    def fl(self, T):
        if self.T_liq <= self.T_sol:
            return np.where(T >= self.T_sol, 1.0, 0.0)
        return np.clip((T - self.T_sol) / (self.T_liq - self.T_sol), 0.0, 1.0)

    def _blend(self, s, l, T):
        fl = self.fl(T)
        return (1.0 - fl) * eval_expr(self.src[s], T) + fl * eval_expr(self.src[l], T)

    def k(self, T):
        return self._blend("k_s", "k_l", T)

    def rho(self, T):
        return self._blend("rho_s", "rho_l", T)

    def c(self, T):
        return self._blend("c_s", "c_l", T)

    def latent(self, T, Tp, L_f, dt):
        """−ρ(T)·L_f·(f_l(T) − f_l(T_prev))/dt with blended ρ (properties.py:375-395)."""
        flc, flp = self.fl(T), self.fl(Tp)
        rho = (1.0 - flc) * eval_expr(self.src["rho_s"], T) + flc * eval_expr(self.src["rho_l"], T)
        return -rho * L_f * (flc - flp) * (1.0 / dt)


def assemble_property_correction(a_trial, T_prev, dt, props, k_bar, a_bar, g):
    """C = project_volume(∇·(k'∇T) − a'∂ₜT) at the trial modes (spectral_ops.py:105-170)."""
    T = reconstruct_volume(a_trial, g)
    kp = props.k(T) - k_bar
    ap = props.rho(T) * props.c(T) - a_bar
    f = correction_source(T, kp, ap, T_prev, float(dt), g.dx, g.dy, g.dz)
    return project_volume(f, g)


# ----------------------------------------------------------------------------------
# Solvers
# ----------------------------------------------------------------------------------


class OracleSolver:
    """FP64 restatement of SpectralSolver.initialize/step/set_state (spectral.py:52-619)."""

    def __init__(self, ctx, *, mixing_omega=None, convergence_tol=None, max_picard_iter=None):
        self.ctx = ctx
        num = ctx.num

        def pick(explicit, attr, default):
            if explicit is not None:
                return explicit
            v = getattr(num, attr, None)
            return v if v is not None else default

        self.mixing_omega = float(pick(mixing_omega, "picard_omega", 0.1))
        self.convergence_tol = float(pick(convergence_tol, "picard_tol", 1e-4))
        self.max_picard_iter = int(pick(max_picard_iter, "max_picard_iter", 30))
        self.track_picard_history = False
        self.picard_history = []
        self._cap_warned = False

    def initialize(self):
        ctx = self.ctx
        geom, num, mat = ctx.geom, ctx.num, ctx.mat
        g = self.grid = OracleGrid(tuple(geom.size), tuple(geom.n))
        n_zyx = (g.nz, g.ny, g.nx)
        self.K, self.KK = precompute_K_KK(float(mat.k), float(mat.rho), float(mat.Cp),
                                          float(num.dt), n_zyx, (g.Lz, g.Ly, g.Lx))
        self.a = np.zeros(n_zyx)
        self.a[0, 0, 0] = float(mat.T0) * math.sqrt(g.Lx * g.Ly * g.Lz)
        self.a_temp = np.empty(n_zyx)
        self.q_evap_old = np.zeros((g.ny, g.nx))
        self.q_evap_buffer = np.zeros((g.ny, g.nx))
        self.n_elements = float(self.a.size)
        model = getattr(mat, "model", None)
        self.prop_corr = model is not None and not model.is_constant
        self.props = OracleProps.from_model(model) if self.prop_corr else None
        fine = getattr(ctx, "fine", None)
        self.fm = (OracleFineMesh(tuple(geom.size), tuple(geom.n), fine.refinement,
                                  tuple(fine.box_size)) if fine is not None else None)
        self.grid_latent = self.fm is None and float(getattr(mat, "L_f", 0.0)) > 0.0
        self.truncation_faces = None
        if self.prop_corr:
            self.kbar = float(mat.k)
            self.abar = float(mat.rho) * float(mat.Cp)
        self.T_prev_full = (np.full(n_zyx, float(mat.T0))
                            if (self.prop_corr or self.grid_latent) else None)
        self.Q_latent_modes_prev = None
        lp = ctx.laser
        self.order = {"gaussian": 2.0, "flat_top": 12.0}.get(
            str(lp.profile).lower(), float(lp.super_gaussian_order))
        self.cell_integrated = bool(lp.cell_integrated)
        return self

    # -- helpers -----------------------------------------------------------------
    def laser_flux(self, ls):
        lp, g = self.ctx.laser, self.grid
        I0 = (float(lp.absorptivity) * float(ls.power)
              / (super_gaussian_area_factor(self.order) * float(lp.r_x) * float(lp.r_y))
              if ls.is_on else 0.0)
        if self.cell_integrated:
            return gaussian_cell_integrated_flux(g.x, g.y, float(ls.x), float(ls.y),
                                                 float(lp.r_x), float(lp.r_y), g.dx, g.dy, I0)
        return super_gaussian_flux(g.x, g.y, float(ls.x), float(ls.y), float(lp.r_x),
                                   float(lp.r_y), self.order, I0)

    def _latent(self, T, Tp):
        mat, num = self.ctx.mat, self.ctx.num
        if self.prop_corr:
            return self.props.latent(T, Tp, float(mat.L_f), float(num.dt))
        return latent_source(T, Tp, float(mat.T_solidus), float(mat.T_liquidus),
                             float(mat.rho), float(mat.L_f), float(num.dt))

    def surface_T(self, a=None):
        return reconstruct_surface(self.a if a is None else a, self.grid)

    def volume_T(self, a=None):
        return reconstruct_volume(self.a if a is None else a, self.grid)

    # -- the step -------------------------------------------------------------------
    def step(self, t, dt):
        ctx, g = self.ctx, self.grid
        mat, num = ctx.mat, ctx.num
        KK, Ct, Cb = self.KK, g.Cp_top, g.Cp_bot
        ls = ctx.laser_path.get_state(t, dt)
        vx, vy = ls.v
        q_las = self.laser_flux(ls)
        P_laser = q_las.sum() * g.dx * g.dy

        self.a *= self.K
        aK = self.a
        a_temp = self.a_temp
        fm = self.fm
        if fm is not None:
            fm.update(ls.x, ls.y)
        a_temp[...] = aK
        if fm is not None:
            # spectral_ops.py:182-195 then :292-306 (history follows the moving box)
            if fm.T_prev is None:
                fm.T_prev = reconstruct_box(a_temp, fm)
                if fm.Q_prev is None:
                    fm.Q_prev = np.zeros(fm.T_prev.shape)
            sy = -float(vy) * dt / fm.d_fine[1]
            sx = -float(vx) * dt / fm.d_fine[0]
            fm.T_prev = shift_3d_xy(fm.T_prev, sy, sx, nearest=True)
            if fm.Q_prev is not None:
                fm.Q_prev = shift_3d_xy(fm.Q_prev, sy, sx, nearest=False)
        if self.track_picard_history:
            self.picard_history = []

        q_evap_shifted = shift_flux(self.q_evap_old, (float(vx) * dt, float(vy) * dt), g.dx, g.dy)
        S_top = g.dct_scale * dct2(q_las - q_evap_shifted)
        h_top, h_bot, T0 = float(mat.h_conv_top), float(mat.h_conv_bottom), float(mat.T0)

        update_modes_etd1(aK, KK, Ct, S_top, a_temp)
        if fm is not None and fm.T_prev is not None and fm.Q_prev is not None:
            add_source_term_modes(a_temp, KK, project_box(fm.Q_prev, fm))
        elif self.grid_latent and self.Q_latent_modes_prev is not None:
            add_source_term_modes(a_temp, KK, self.Q_latent_modes_prev)
        if h_top > 0:
            Ts0 = reconstruct_surface(a_temp, g)
            S = g.dct_scale * dct2(-h_top * (Ts0 - T0))
            add_bottom_surface_source(a_temp, KK, Ct, S)
        if h_bot > 0:
            Tb = reconstruct_bottom(a_temp, g)
            S_bot = g.dct_scale * dct2(-h_bot * (Tb - T0))
            add_bottom_surface_source(a_temp, KK, Cb, S_bot)

        S_las = g.dct_scale * dct2(q_las)
        rescale = self.prop_corr
        S_bot_raw = None
        if h_bot > 0:
            Tb = reconstruct_bottom(a_temp, g)
            q = -h_bot * (Tb - T0)
            if rescale:
                q = q * (self.kbar / self.props.k(Tb))
            S_bot_raw = g.dct_scale * dct2(q)

        iter_k, rel, conv = -1, float("nan"), False
        Q_modes = Q_raw = None
        for iter_k in range(self.max_picard_iter):
            a_old = a_temp.copy()
            Ts = reconstruct_surface(a_old, g)
            self.q_evap_buffer = evaporation_flux(Ts, float(mat.Pa), float(mat.T_boil),
                                                  float(mat.DeltaH_LV), float(mat.R_v),
                                                  float(mat.T_liquidus))
            q_conv = h_top * (Ts - T0) if h_top > 0 else 0.0
            if rescale:
                q_top = (q_las - self.q_evap_buffer - q_conv) * (self.kbar / self.props.k(Ts))
                S_top_raw = g.dct_scale * dct2(q_top)
            else:
                S_top_raw = S_las - g.dct_scale * dct2(self.q_evap_buffer)
                if h_top > 0:
                    S_top_raw = S_top_raw - g.dct_scale * dct2(q_conv)
            Q_modes = Q_raw = None
            if fm is not None and fm.T_prev is not None:
                Q_raw = self._latent(reconstruct_box(a_old, fm), fm.T_prev)
                Q_modes = project_box(Q_raw, fm)
            elif self.grid_latent:
                T_full = reconstruct_volume(a_old, g)
                Q_modes = project_volume(self._latent(T_full, self.T_prev_full), g)
            update_modes_etd1(aK, KK, Ct, S_top_raw, a_temp)
            if Q_modes is not None:
                add_source_term_modes(a_temp, KK, Q_modes)
            if S_bot_raw is not None:
                add_bottom_surface_source(a_temp, KK, Cb, S_bot_raw)
            if self.prop_corr:
                C = assemble_property_correction(a_temp, self.T_prev_full, num.dt, self.props,
                                                 self.kbar, self.abar, g)
                add_source_term_modes(a_temp, KK, C)
            r = a_temp - a_old
            np.multiply(r, self.mixing_omega, out=a_temp)
            a_temp += a_old
            rms_diff = math.sqrt(float(np.vdot(r, r)) / self.n_elements)
            rms_old = math.sqrt(float(np.vdot(a_old, a_old)) / self.n_elements)
            rel = rms_diff / max(rms_old, 1e-9)
            if self.track_picard_history:
                rho_k = None
                if iter_k > 0 and self.picard_history:
                    prev = self.picard_history[-1]["rms_diff"]
                    if prev > 0:
                        rho_k = rms_diff / prev
                self.picard_history.append({"iter": iter_k, "rms_diff": rms_diff,
                                            "rho": rho_k, "true_rel_err": rel})
            if rel < self.convergence_tol:
                conv = True
                break
        else:
            self._cap_warned = True

        T_temp = reconstruct_surface(a_temp, g)
        if self.grid_latent:
            self.Q_latent_modes_prev = Q_modes
        self.q_evap_old = self.q_evap_buffer.copy()
        self.a[...] = a_temp
        if self.prop_corr or self.grid_latent:
            self.T_prev_full = reconstruct_volume(self.a, g)
        if fm is not None:
            # spectral_ops.py:198-207 and spectral.py:524-534
            fm.T_prev = reconstruct_box(a_temp, fm)
            if fm.Q_prev is None:
                fm.Q_prev = np.zeros(fm.T_prev.shape)
            if Q_raw is not None:
                fm.Q_prev = Q_raw.copy()
                if self.truncation_faces is None:
                    hit = box_face_peaks(Q_raw, fm)
                    if hit:
                        self.truncation_faces = hit
        return self, {"T_surface_max": float(T_temp.max()), "P_laser": float(P_laser),
                      "n_evap_iter": iter_k + 1, "picard_converged": conv,
                      "picard_rel_err": rel}

    def set_state(self, T):
        g = self.grid
        self.a = dct2(np.asarray(T, dtype=np.float64)) * math.sqrt(g.dx * g.dy * g.dz)


class OracleLinearSolver:
    """Laser-only ETD1 stepper (spectral_cpu_linear.py:61-162)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def initialize(self):
        ctx = self.ctx
        g = self.grid = OracleGrid(tuple(ctx.geom.size), tuple(ctx.geom.n))
        mat, num = ctx.mat, ctx.num
        self.K, self.KK = precompute_K_KK(float(mat.k), float(mat.rho), float(mat.Cp),
                                          float(num.dt), (g.nz, g.ny, g.nx), (g.Lz, g.Ly, g.Lx))
        self.a = np.zeros((g.nz, g.ny, g.nx))
        self.a[0, 0, 0] = float(mat.T0) * np.sqrt(g.Lx * g.Ly * g.Lz)
        return self

    def step(self, t, dt):
        ctx, g = self.ctx, self.grid
        lp = ctx.laser
        ls = ctx.laser_path.get_state(t, dt)
        coef = (float(lp.absorptivity) * 2.0 * float(ls.power) / (np.pi * float(lp.radius) ** 2)
                if ls.is_on else 0.0)
        r = float(lp.radius)
        q = super_gaussian_flux(g.x, g.y, float(ls.x), float(ls.y), r, r, 2.0, coef)
        self.a *= self.K
        S = g.dct_scale * dct2(q)
        out = np.empty_like(self.a)
        update_modes_etd1(self.a, self.KK, g.Cp_top, S, out)
        self.a = out
        return self, {"P_laser": float(q.sum() * g.dx * g.dy)}


def reconstruct_at_points(a, g, coords):
    """T at arbitrary points (N, 3), coordinates clipped to the domain
    (spectral_helpers.py:209-255): Σ a[p,n,m]·Bz[p,i]·By[n,i]·Bx[m,i]."""
    c = np.asarray(coords, dtype=np.float64)
    x = np.clip(c[:, 0], 0.0, g.Lx)
    y = np.clip(c[:, 1], 0.0, g.Ly)
    z = np.clip(c[:, 2], 0.0, g.Lz)
    B = [g.C[a_][:, None] * np.cos(np.pi * np.arange(n)[:, None] * v[None, :] / L)
         for a_, (n, L, v) in enumerate(((g.nx, g.Lx, x), (g.ny, g.Ly, y), (g.nz, g.Lz, z)))]
    return np.einsum("pnm,pi,ni,mi->i", a, B[2], B[1], B[0], optimize=True)


def reconstruct_temperature_DCT(a, g):
    """Node-centred (nx+1, ny+1, nz+1) field via an unnormalised DCT-I (spectral_helpers.py:147-207)."""
    nz, ny, nx = a.shape
    wx, wy, wz = (c.copy() for c in g.C)
    wx[1:] *= 0.5
    wy[1:] *= 0.5
    wz[1:] *= 0.5
    b = a * wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
    padded = np.zeros((nz + 1, ny + 1, nx + 1))
    padded[:nz, :ny, :nx] = b
    T = scipy.fft.dctn(padded, type=1, norm=None, axes=(0, 1, 2), workers=FFT_WORKERS)
    return np.ascontiguousarray(T.transpose(2, 1, 0))


def trapezoidal_weights(c):
    w = np.zeros_like(c)
    d = np.diff(c)
    w[:-1] += d * 0.5
    w[1:] += d * 0.5
    return w


def compute_L2_structured(Ta, Tb, x, y, z):
    """Trapezoid-weighted L2 / relative L2 / Linf on a [z,y,x] grid (compute_L2_error.py:491-531)."""
    err = Ta - Tb
    W = (trapezoidal_weights(z)[:, None, None] * trapezoidal_weights(y)[None, :, None]
         * trapezoidal_weights(x)[None, None, :])
    L2 = np.sqrt(np.sum(W * err ** 2))
    ref = np.sum(W * Ta ** 2)
    return {"L2_abs": float(L2), "L2_rel": float(L2 / np.sqrt(ref)) if ref > 0 else np.inf,
            "Linf": float(np.max(np.abs(err)))}


def modal_rel_l2(a_test, a_ref):
    """Relative L2 of two modal states; equals the field rel-L2 by Parseval."""
    a_test = np.asarray(a_test, dtype=np.float64)
    a_ref = np.asarray(a_ref, dtype=np.float64)
    return float(np.linalg.norm(a_test - a_ref) / np.linalg.norm(a_ref))


def modal_rel_l2_excursion(a_test, a_ref, mean0):
    """Relative L2 of (T − T0): T0 lives in mode (0,0,0) as ``mean0 = T0·√V``, which
    cancels in the difference and is subtracted from the reference norm."""
    a_test = np.asarray(a_test, dtype=np.float64)
    a_ref = np.asarray(a_ref, dtype=np.float64)
    diff = np.linalg.norm(a_test - a_ref)
    a_ref_ex = a_ref.copy()
    a_ref_ex[0, 0, 0] -= mean0
    base = np.linalg.norm(a_ref_ex)
    return float(diff / base) if base > 0 else float("inf")
