"""Problem set-up: initial conditions of the reference's drivers, restated.

* ``riemann_prim`` follows ``IC_for_riemann_2d`` (reference
  ``src/2Dproblem.cpp:345-385``): primitive states by quadrant for interior
  cells, zeros in the ghost layers, with the quadrant tests written against
  ``discon/dx`` on both axes exactly as the reference does.
* ``riemann_states(seed)`` draws the four quadrant states of the seeded
  four-quadrant Riemann workload with ``std::mt19937_64(seed)`` and
  ``std::uniform_real_distribution<double>`` semantics (libstdc++
  ``generate_canonical``), in the order zone 1..4, each (rho, u, v, p), with
  rho in [0.1, 1.5], u, v in [-0.8, 0.8], p in [0.02, 1.5].
"""
import numpy as np

_MASK64 = (1 << 64) - 1


class MT19937_64:
    """The 64-bit Mersenne Twister (std::mt19937_64)."""

    def __init__(self, seed=5489):
        self.mt = [0] * 312
        self.mt[0] = seed & _MASK64
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & _MASK64
        self.idx = 312

    def _twist(self):
        mt = self.mt
        for i in range(312):
            x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + 156) % 312] ^ xa
        self.idx = 0

    def __call__(self):
        if self.idx >= 312:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & _MASK64

    def uniform(self, a, b):
        # generate_canonical<double, 53> with a 64-bit engine: one draw
        r = float(self()) / 18446744073709551616.0
        if r >= 1.0:
            r = np.nextafter(1.0, 0.0)
        return r * (b - a) + a


RIEMANN_RANGES = ((0.1, 1.5), (-0.8, 0.8), (-0.8, 0.8), (0.02, 1.5))


def riemann_states(seed):
    """Four quadrant primitive states (zone1..zone4, each rho,u,v,p)."""
    g = MT19937_64(seed)
    return np.array([[g.uniform(*RIEMANN_RANGES[k]) for k in range(4)] for _ in range(4)])


def riemann_prim(p, zones, discon=(0.5, 0.5)):
    """Primitive field [4, nx, ny] of IC_for_riemann_2d (ghosts zero)."""
    G, nx, ny = p.ghost, p.nx, p.ny
    dx = (p.right - p.left) / p.nodex
    zones = np.asarray(zones, dtype=np.float64)
    i = np.arange(G, nx - G)
    j = np.arange(G, ny - G)
    xi = ((i - G + 1) < discon[0] / dx)[:, None]
    yj = ((j - G + 1) < discon[1] / dx)[None, :]
    z = np.where(xi & yj, 0, np.where(~xi & yj, 1, np.where(~xi & ~yj, 2, 3)))
    prim = np.zeros((4, nx, ny))
    prim[:, G:nx - G, G:ny - G] = zones[z].transpose(2, 0, 1)
    return prim


def prim_to_convar(prim, gamma):
    """Primvar_to_convar_2D (reference src/basic_function.cpp:860-866) on a
    [4, ...] array; ghost cells with rho=0 give NaN energy as in the
    reference."""
    r, u, v, pr = prim
    with np.errstate(invalid="ignore", divide="ignore"):
        e = r * (pr / r / (gamma - 1) + 0.5 * (u * u + v * v))
    return np.stack([r, r * u, r * v, e])


# RM2 configuration used by PlanarShock (reference src/2Dproblem.cpp:86-92)
PLANAR_SHOCK_ZONES = np.array([
    [0.138, 1.206, 1.206, 0.029],
    [0.5323, 0.0, 1.206, 0.3],
    [1.5, 0.0, 0.0, 1.5],
    [0.5323, 1.206, 0.0, 0.3],
])


# ------------------------------------------------------------------------ 1D
def _fill_ghosts_1d(prim, G):
    prim[:, :G] = prim[:, G:G + 1]
    prim[:, -G:] = prim[:, -G - 1:-G]
    return prim


def sod_prim(p):
    """ICforSodTube (reference src/1Dproblem.cpp:151-173): (1, 0, 1) where
    (i - ghost + 1) * dx <= 0.5, else (0.125, 0, 0.1); [3, nx].  The
    reference leaves the left ghost layer unset; here it repeats the first
    interior cell (ghosts are overwritten by the boundary fill before use)."""
    G, nx, dx = p.ghost, p.nx, p.dx
    prim = np.zeros((3, nx))
    i = np.arange(G, nx)
    left = (i - G + 1).astype(np.float64) * dx <= 0.5
    prim[:, G:] = np.where(left[None, :], np.array([[1.0], [0.0], [1.0]]), np.array([[0.125], [0.0], [0.1]]))
    return _fill_ghosts_1d(prim, G)


def blastwave_prim(p):
    """ICforBlastwave (reference src/1Dproblem.cpp:323-345): p = 1000 where
    (i - ghost + 1) * dx <= 0.1, p = 100 where >= 0.9, else 0.01; rho 1, u 0."""
    G, nx, dx = p.ghost, p.nx, p.dx
    prim = np.zeros((3, nx))
    i = np.arange(G, nx - G)
    x = (i - G + 1).astype(np.float64) * dx
    prim[0, G:nx - G] = 1.0
    prim[2, G:nx - G] = np.where(x <= 0.1, 1000.0, np.where(x >= 0.9, 100.0, 0.01))
    return _fill_ghosts_1d(prim, G)


def shu_osher_prim(p):
    """ICforShuOsher (reference src/1Dproblem.cpp:493-515), cell-averaged
    density wave right of x = 1."""
    G, nx, dx = p.ghost, p.nx, p.dx
    prim = np.zeros((3, nx))
    for i in range(G, nx):
        x = (i - G + 0.5) * dx
        if x <= 1.0:
            prim[:, i] = (3.85714, 2.629369, 10.33333)
        else:
            prim[:, i] = (1 - (0.2 / (5 * dx)) * (np.cos(5 * (x + 0.5 * dx)) - np.cos(5 * (x - 0.5 * dx))), 0.0, 1.0)
    return _fill_ghosts_1d(prim, G)


def sinwave_prim_1d(p):
    """ICfor_sinwave (reference src/1Dproblem.cpp:692-706): cell averages of
    rho = 1 + 0.2 sin(pi x), u = 1, p = 1 on the interior."""
    G, nx, dx = p.ghost, p.nx, p.dx
    pi = 3.14159265358979323846
    prim = np.zeros((3, nx))
    for i in range(G, nx - G):
        prim[0, i] = 1.0 - 0.2 / pi / dx * (np.cos(pi * (i + 1 - G) * dx) - np.cos(pi * (i - G) * dx))
        prim[1, i] = 1.0
        prim[2, i] = 1.0
    return _fill_ghosts_1d(prim, G)


def sinwave_exact_1d(p):
    """Exact cell-averaged density of the sine wave at any t that is a whole
    number of periods (error_for_sinwave, reference src/1Dproblem.cpp:708-735)."""
    return sinwave_prim_1d(p)[0, p.ghost:p.nx - p.ghost]


# ------------------------------------------------------------- 2D sine wave
def sinwave2d_rho(p):
    """Cell averages of rho = 1 + 0.2 sin(pi (x + y)) on every cell of an
    [nx, ny] block (ICfor_sinwave_2d / error_for_sinwave_2d, reference
    src/2Dproblem.cpp:1529-1601: the four corner sines of each cell in the
    reference's operation order; the ghost layers get the periodic
    continuation, which the boundary fills overwrite anyway)."""
    G = p.ghost
    dx = (p.right - p.left) / p.nodex
    dy = (p.up - p.down) / p.nodey
    pi = 3.14159265358979323846
    i = np.arange(p.nx, dtype=np.float64)[:, None]
    j = np.arange(p.ny, dtype=np.float64)[None, :]
    xl, xr = (i - G) * dx, (i + 1 - G) * dx
    yl, yr = (j - G) * dy, (j + 1 - G) * dy
    k1 = np.sin(pi * (xr + yr))
    k2 = np.sin(pi * (xr + yl))
    k3 = np.sin(pi * (xl + yr))
    k4 = np.sin(pi * (xl + yl))
    return 1.0 - 0.2 / pi / pi / dx / dy * ((k1 - k2) - (k3 - k4))


def sinwave2d_prim(p):
    """Primitive field [4, nx, ny] of ICfor_sinwave_2d: the density wave,
    u = v = 1, p = 1."""
    prim = np.ones((4, p.nx, p.ny))
    prim[0] = sinwave2d_rho(p)
    return prim


# ------------------------------------------------------------ isentropic vortex
def vortex_prim_point(x, y, gamma=1.4, eps=5.0, xc=5.0, yc=5.0, free=(1.0, 1.0, 1.0, 1.0)):
    """Point values (rho, u, v, p) of the isentropic vortex of strength eps
    centred at (xc, yc) in the free stream (rho, u, v, p) = free (with
    p/rho = 1): du = -eps/(2 pi) e^((1-r^2)/2) (y-yc), dv = eps/(2 pi)
    e^((1-r^2)/2) (x-xc), T = 1 - (gamma-1) eps^2/(8 gamma pi^2) e^(1-r^2),
    rho = T^(1/(gamma-1)), p = rho T."""
    r2 = (x - xc) ** 2 + (y - yc) ** 2
    a = eps / (2 * np.pi) * np.exp(0.5 * (1.0 - r2))
    T = 1.0 - (gamma - 1) * eps * eps / (8 * gamma * np.pi ** 2) * np.exp(1.0 - r2)
    rho = free[0] * T ** (1.0 / (gamma - 1))
    return np.stack([rho, free[1] - a * (y - yc), free[2] + a * (x - xc), rho * T])


def vortex_convar(p, eps=5.0, quad=5, shift=(0.0, 0.0)):
    """Cell averages of the conservative variables [4, nx, ny] (ghost cells
    included) by quad x quad Gauss-Legendre points per cell; the vortex
    centre moved by `shift` and wrapped into the periodic square (the exact
    solution at time t is the initial one shifted by t (u, v) = (t, t))."""
    G, nx, ny = p.ghost, p.nx, p.ny
    dx = (p.right - p.left) / p.nodex
    dy = (p.up - p.down) / p.nodey
    Lx, Ly = p.right - p.left, p.up - p.down
    xg, wg = np.polynomial.legendre.leggauss(quad)
    xc0 = np.mod(5.0 + shift[0], Lx)
    yc0 = np.mod(5.0 + shift[1], Ly)
    out = np.zeros((4, nx, ny))
    cx = p.left + (np.arange(nx) - G + 0.5) * dx
    cy = p.down + (np.arange(ny) - G + 0.5) * dy
    for a, wa in zip(xg, wg):
        for b, wb in zip(xg, wg):
            X = (cx + 0.5 * dx * a)[:, None] * np.ones((1, ny))
            Y = np.ones((nx, 1)) * (cy + 0.5 * dy * b)[None, :]
            # nearest periodic image of the centre
            ddx = np.mod(X - xc0 + 0.5 * Lx, Lx) - 0.5 * Lx
            ddy = np.mod(Y - yc0 + 0.5 * Ly, Ly) - 0.5 * Ly
            pr = vortex_prim_point(ddx, ddy, p.gamma, eps, 0.0, 0.0)
            out += 0.25 * wa * wb * prim_to_convar(pr, p.gamma)
    return out


def convar_to_prim(c, gamma):
    """Convar_to_primvar_2D (reference src/basic_function.cpp:811-817)."""
    r, mu, mv, E = c
    return np.stack([r, mu / r, mv / r, (gamma - 1.0) * (E - 0.5 * mu * mu / r - 0.5 * mv * mv / r)])


# --------------------------------------------------------------------- 3D
def prim_to_convar_3d(prim, gamma):
    """Primvar_to_convar_3D (basic_function.cpp:859-866) on [5, ...] arrays."""
    r, u, v, w, p = prim
    return np.stack([r, r * u, r * v, r * w, r * (p / r / (gamma - 1) + 0.5 * (u * u + v * v + w * w))])


def convar_to_prim_3d(c, gamma):
    r = c[0]
    u, v, w = c[1] / r, c[2] / r, c[3] / r
    p = (gamma - 1.0) * (c[4] - 0.5 * c[1] * c[1] / r - 0.5 * c[2] * c[2] / r - 0.5 * c[3] * c[3] / r)
    return np.stack([r, u, v, w, p])


def cubic_sod_prim(p, inner=(1.0, 0.0, 0.0, 0.0, 1.0), outer=(0.125, 0.0, 0.0, 0.0, 0.1)):
    """ICfor_cubic_sod (reference src/3Dproblem.cpp:174-223): the cube of
    cells whose centres lie within 0.2 of the domain centre in every
    direction holds `inner`, the rest `outer`; ghost cells copy the nearest
    interior cell (they are overwritten by the first boundary fill)."""
    G = p.ghost
    xs = p.left + (np.arange(p.nx) + 0.5 - G) * p.dx
    ys = p.down + (np.arange(p.ny) + 0.5 - G) * p.dy
    zs = p.back + (np.arange(p.nz) + 0.5 - G) * p.dz
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    box = ((np.abs(X - 0.5 * (p.right + p.left)) < 0.2 * (p.right - p.left)) &
           (np.abs(Y - 0.5 * (p.up + p.down)) < 0.2 * (p.up - p.down)) &
           (np.abs(Z - 0.5 * (p.front + p.back)) < 0.2 * (p.front - p.back)))
    prim = np.empty((5, p.nx, p.ny, p.nz))
    for v in range(5):
        prim[v] = np.where(box, inner[v], outer[v])
    return prim


def tgv_prim_point(x, y, z, mach=0.1, rho0=1.0, u0=1.0, gamma=1.4):
    """Taylor-Green vortex at t=0: u = U0 sin x cos y cos z, v = -U0 cos x
    sin y cos z, w = 0, p = p0 + rho0 U0^2 / 16 (cos 2x + cos 2y)(cos 2z + 2)
    with p0 = rho0 U0^2 / (gamma Ma^2), rho = rho0."""
    p0 = rho0 * u0 * u0 / (gamma * mach * mach)
    u = u0 * np.sin(x) * np.cos(y) * np.cos(z)
    v = -u0 * np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    pr = p0 + rho0 * u0 * u0 / 16.0 * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    return np.stack([np.full_like(u, rho0), u, v, w, pr])


def tgv_convar(p, mach=0.1, quad=3):
    """Cell averages of the conservative variables of the Taylor-Green vortex
    (Gauss-Legendre quadrature with `quad` points per direction), ghost
    cells included (periodic)."""
    G = p.ghost
    gx, gw = np.polynomial.legendre.leggauss(quad)
    xs = p.left + (np.arange(p.nx) + 0.5 - G) * p.dx
    ys = p.down + (np.arange(p.ny) + 0.5 - G) * p.dy
    zs = p.back + (np.arange(p.nz) + 0.5 - G) * p.dz
    out = np.zeros((5, p.nx, p.ny, p.nz))
    for a, wa in zip(gx, gw):
        for b, wb in zip(gx, gw):
            for c, wc in zip(gx, gw):
                X, Y, Z = np.meshgrid(xs + 0.5 * a * p.dx, ys + 0.5 * b * p.dy, zs + 0.5 * c * p.dz, indexing="ij")
                out += (wa * wb * wc / 8.0) * prim_to_convar_3d(tgv_prim_point(X, Y, Z, mach, gamma=p.gamma), p.gamma)
    return out


def kinetic_energy(convar, p):
    """Mean kinetic energy 0.5 rho |u|^2 over the interior cells."""
    G = p.ghost
    c = convar[:, G:G + p.nodex, G:G + p.nodey, G:G + p.nodez]
    return float(np.mean(0.5 * (c[1] ** 2 + c[2] ** 2 + c[3] ** 2) / c[0]))
