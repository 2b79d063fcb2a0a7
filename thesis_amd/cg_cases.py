"""The reference's two-phase test cases (D2Q9): initial colour fields and
the host-side validators that read the results.

Initial shapes (initCGBubble, S/GpuInit.cu:323-534): 1 steady bubble,
2 two layers (Couette / two-phase Poiseuille), 3 square relaxing to a circle,
4 two touching bubbles coalescing, 5 elliptic bubble oscillating, 6 heavy
layer over a light one with a cosine perturbation (Rayleigh-Taylor).  Each
returns the red density field (ny, nx) of a unit-density start; blue is one
minus it.

Validators (S/Multiphase.cu:1104-1340) read colour densities (ny, nx), or
(nz, ny, nx) for the 3D cases, where they look at the middle plane; the
"middle" row, column or plane is index m/2 for even m and (m+1)/2 for odd m,
as the reference takes it.

Coordinates: the reference reads node coordinates from the mesh; here x
runs over one period, x = i / nx (the box is periodic in x), and y uses the
same spacing, y = j / nx, unless a case says otherwise.
"""
import math

import numpy as np


def _mid(m):
    return m // 2 if m % 2 == 0 else (m + 1) // 2


def _phase(r_rho, b_rho):
    r_rho = np.asarray(r_rho, float)
    b_rho = np.asarray(b_rho, float)
    return (r_rho - b_rho) / (r_rho + b_rho)


# ---- initial shapes -------------------------------------------------------
def bubble(nx, ny, radius):
    """Case 1: red disc of `radius` (in units of the box width) at the centre."""
    y, x = np.mgrid[0:ny, 0:nx]
    xs, ys = x / (nx - 1), y / (ny - 1)
    return (np.hypot(xs - 0.5, ys - 0.5) <= radius).astype(float)


def layers(nx, ny):
    """Case 2: red above y = 1/2 of the channel height (y = j / (ny - 1))."""
    y = np.arange(ny)[:, None] / (ny - 1)
    return np.where(y > 0.5, 1.0, 0.0) * np.ones((1, nx))


def square(nx, ny):
    """Case 3: red square of half the box, by node index."""
    y, x = np.mgrid[0:ny, 0:nx]
    inside = (x < 0.75 * nx) & (x > 0.25 * nx) & (y < 0.75 * ny) & (y > 0.25 * ny)
    return inside.astype(float)


def two_bubbles(nx, ny, radius):
    """Case 4: two red discs of `radius` touching at the centre, one above the other."""
    y, x = np.mgrid[0:ny, 0:nx]
    xs, ys = x / (nx - 1), y / (ny - 1)
    a = np.hypot(xs - 0.5, ys - 0.5 + radius) <= radius
    b = np.hypot(xs - 0.5, ys - 0.5 - radius) <= radius
    return (a | b).astype(float)


def ellipse(nx, ny):
    """Case 5: red ellipse of semi-axes nx/8 and 3 ny/16 at the centre, by node index."""
    y, x = np.mgrid[0:ny, 0:nx]
    inside = ((x - 0.5 * nx) / (0.125 * nx)) ** 2 + ((y - 0.5 * ny) / (0.1875 * ny)) ** 2 <= 1
    return inside.astype(float)


def rayleigh_taylor(nx, ny, amplitude=0.1):
    """Case 6: red (heavy) above y0 + amplitude cos(2 pi x), y0 the middle of
    the box (the reference's 1 x 4 box puts it at y = 2)."""
    y, x = np.mgrid[0:ny, 0:nx]
    y0 = 0.5 * (ny - 1) / nx
    return (y / nx > y0 + amplitude * np.cos(2 * np.pi * x / nx)).astype(float)


# ---- validators -------------------------------------------------------------
def analytical_couette(kappa, y, u_wall):
    """Two-layer Couette profile (analyticalCouette, S/Multiphase.cu:1136-1160):
    y the node heights of one column scaled to [0, 1], kappa = mu_r / mu_b
    (red on top); rows from the middle up take the upper layer's line."""
    y = np.asarray(y, float)
    m = len(y)
    a = np.empty(m)
    j0 = _mid(m)
    a[j0:] = (2.0 * y[j0:] / (kappa + 1.0) + (kappa - 1.0) / (kappa + 1.0)) * u_wall
    a[:j0] = (2.0 * kappa * y[:j0] / (kappa + 1.0)) * u_wall
    return a


def analytical_poiseuille(y, r_density, b_density, r_visc, b_visc, g):
    """Two-layer force-driven channel profile (analyticalPoiseuille,
    S/Multiphase.cu:1303-1340), y the node heights of one column."""
    y = np.asarray(y, float)
    m = len(y)
    r_mu, b_mu = r_density * r_visc, b_density * b_visc
    a = np.empty(m)
    j0 = _mid(m)
    yu = y[j0:]
    a[j0:] = (-g * yu * yu / (2.0 * r_mu) + g * ((3.0 * b_mu + r_mu) * yu / (4.0 * r_mu * (r_mu + b_mu)))
              + g / (2.0 * r_mu) - g * ((3.0 * b_mu + r_mu) / (4.0 * r_mu * (r_mu + b_mu))))
    yl = y[:j0]
    a[:j0] = -g * yl * yl / (2.0 * b_mu) + g * ((3.0 * b_mu + r_mu) * yl) / (4.0 * b_mu * (r_mu + b_mu))
    return a


def mid_row_red_width(r_rho, b_rho):
    """Sum of the positive phase field along the middle row (of the middle
    plane in 3D): the red chord through the centre, in nodes
    (deformingBubbleValid / validateCoalescenceCase)."""
    phi = _phase(r_rho, b_rho)
    if phi.ndim == 3:
        phi = phi[_mid(phi.shape[0])]
    row = phi[_mid(phi.shape[0])]
    return float(row[row > 0.0].sum())


def deforming_radius(r_rho, b_rho):
    """2D case 3 / 3D case 2 (deformingBubbleValid, S/Multiphase.cu:1162-1192):
    measured and predicted radius in nodes; the square of side n/2 becomes
    the disc of the same area, radius n / (2 sqrt(pi)), the cube the sphere
    of the same volume, radius (n/2) / (4 pi / 3)^(1/3)."""
    a = np.asarray(r_rho)
    nx = a.shape[-1]
    want = nx / 2.0 / (4.0 / 3.0 * math.pi) ** (1.0 / 3.0) if a.ndim == 3 else nx / (2.0 * math.sqrt(math.pi))
    return mid_row_red_width(r_rho, b_rho) / 2.0, want


def coalescence_radius(r_rho, b_rho, radius):
    """2D case 4 / 3D case 3 (validateCoalescenceCase, S/Multiphase.cu:1104-1134):
    measured radius in units of the box width and the predicted radius of the
    merged drop, sqrt(2) R for two discs, 2^(1/3) R for two spheres."""
    a = np.asarray(r_rho)
    nx = a.shape[-1]
    want = radius * (2.0 ** (1.0 / 3.0) if a.ndim == 3 else math.sqrt(2.0))
    return mid_row_red_width(r_rho, b_rho) / (2.0 * nx), want


def _crossing(phi_a, phi_b, y_a, y_b):
    slope = (phi_a - phi_b) / (y_a - y_b)
    return -(phi_a - slope * y_a) / slope


def max_y_oscillating(r_rho, b_rho, y=None):
    """Case 5 (getMaxYOscilating, S/Multiphase.cu:1211-1239): height of the
    top of the red region on the middle column, the zero of the phase field
    interpolated between the first blue node above the middle and the node
    below it."""
    phi = _phase(r_rho, b_rho)
    m, n = phi.shape
    i = _mid(n)
    yy = np.arange(m, dtype=float) if y is None else np.asarray(y, float)
    j = _mid(m)
    while j < m - 1 and phi[j, i] >= 0.0:
        j += 1
    return _crossing(phi[j, i], phi[j - 1, i], yy[j], yy[j - 1])


def min_y_rt(r_rho, b_rho, y=None):
    """Case 6 (getMinYRT, S/Multiphase.cu:1241-1269): height of the bottom of
    the red (heavy) region on the middle column, scanning down from the middle."""
    phi = _phase(r_rho, b_rho)
    m, n = phi.shape
    i = _mid(n)
    yy = np.arange(m, dtype=float) if y is None else np.asarray(y, float)
    j = _mid(m)
    while j > 0 and phi[j, i] >= 0.0:
        j -= 1
    return _crossing(phi[j, i], phi[j + 1, i], yy[j], yy[j + 1])


def oscillation_frequency(extremes, sigma, r_density, b_density, radius, start=500, every=1):
    """Case 5 (validateOscilating, S/Multiphase.cu:1271-1301): from the trace
    of max_y_oscillating, the first maximum and second minimum after `start`
    steps give the numerical angular frequency pi / (t_min2 - t_max1); the
    theoretical one is sqrt(6 sigma / ((rho_r + rho_b) R^3)), the 2D mode-2
    frequency of Lamb.  The reference samples every step; a trace sampled
    every `every` steps is read the same way with its times scaled (sampling
    steps apart keeps the step-to-step jitter of the interpolated interface
    out of the extremum search).  Returns (theoretical, numerical, error %)."""
    e = np.asarray(extremes, float)
    maxima, minima = [], []
    rising = False
    for i in range(max(1, -(-start // every)), len(e) - 1):
        d = (e[i + 1] - e[i - 1]) / 2.0
        if rising and d < 0.0:
            maxima.append(i)
            rising = False
        elif not rising and d > 0.0:
            minima.append(i)
            rising = True
    theo = math.sqrt(6.0 * sigma / ((r_density + b_density) * radius ** 3))
    if not maxima or len(minima) < 2:
        return theo, float("nan"), float("nan")
    num = math.pi / ((minima[1] - maxima[0]) * every)
    return theo, num, abs(theo - num) / theo * 100.0
