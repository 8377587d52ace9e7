"""Run monitor with the reference's interface (io_utils/monitor.py:14-260):
records density range, peak speed, mass (sum of rho), minimum air-film
thickness and, in 2D, the left / right contact angles from the density
gradient at the contact columns.

It reads the solver's observables (rho, u, solid) once per record() and
evaluates them on the host, as the reference does.
"""
import numpy as np


class Monitor:
    """History of per-record quantities; see record() and summary()."""

    KEYS = ("step", "t", "rho_min", "rho_max", "u_max", "mass", "film_thickness",
            "contact_angle_left", "contact_angle_right")

    def __init__(self, config, log_file=None):
        self.config = config
        self.log_file = log_file
        self.history = {k: [] for k in self.KEYS}
        self.initial_mass = None

    # -- recording -------------------------------------------------------------
    def record(self, solver, step=None):
        rho = solver.rho.cpu().numpy()
        u = solver.u.cpu().numpy()
        solid = solver.solid.cpu().numpy()
        speed = np.sqrt((u ** 2).sum(axis=0))
        mass = float(rho.sum())
        if self.initial_mass is None:
            self.initial_mass = mass
        film = self.compute_film_thickness(rho, solid)
        ca_l, ca_r = self.compute_contact_angle(rho, solid)
        h = self.history
        h["step"].append(step or solver.step_count)
        h["t"].append(solver.t)
        h["rho_min"].append(float(rho.min()))
        h["rho_max"].append(float(rho.max()))
        h["u_max"].append(float(speed.max()))
        h["mass"].append(mass)
        h["film_thickness"].append(film)
        h["contact_angle_left"].append(ca_l)
        h["contact_angle_right"].append(ca_r)

    # -- air film -------------------------------------------------------------
    def _threshold(self):
        return (self.config.rho_l + self.config.rho_g) / 2

    def _gaps(self, rho, solid):
        """Per column (last axis = wall normal): substrate top index (-1 if
        none), first liquid index above it (rho > threshold), and whether the
        column holds liquid above the substrate."""
        thr = self._threshold()
        n = rho.shape[-1]
        levels = np.arange(n)
        top = np.where(solid, levels, -1).max(axis=-1)
        above = levels > top[..., None]
        liquid = above & (rho > thr)
        first = np.where(liquid, levels, n).min(axis=-1)
        return top, first, liquid.any(axis=-1), n

    def compute_film_thickness(self, rho, solid):
        """Smallest gap (liquid bottom - substrate top) over the columns with
        liquid, times config.dx; inf without liquid."""
        top, first, has, n = self._gaps(rho, solid)
        if not has.any():
            return float("inf")
        gap = np.where(has, first - top, n)
        g = gap[has].min()
        return g * self.config.dx if g < n else float("inf")

    def compute_film_thickness_map(self, rho, solid):
        """Gap per column in lattice units (NaN where a column holds no
        liquid)."""
        top, first, has, _ = self._gaps(rho, solid)
        gap = np.where(has, first - top, 0).astype(float)
        gap[~has] = np.nan
        return gap

    # -- contact angle (2D) ------------------------------------------------------
    def compute_contact_angle(self, rho, solid):
        """(left, right) angles in degrees from the density gradient at the
        outermost columns with liquid on the first layer above the substrate;
        (nan, nan) in 3D or without contact."""
        if rho.ndim != 2:
            return float("nan"), float("nan")
        nx, ny = rho.shape
        thr = self._threshold()
        top = np.where(solid, np.arange(ny), -1).max(axis=1)
        wet = np.array([top[i] + 1 < ny and rho[i, top[i] + 1] > thr for i in range(nx)])
        if not wet.any():
            return float("nan"), float("nan")
        cols = np.nonzero(wet)[0]
        gx = np.zeros_like(rho)
        gy = np.zeros_like(rho)
        gx[1:-1, :] = (rho[2:, :] - rho[:-2, :]) / 2.0
        gy[:, 1:-1] = (rho[:, 2:] - rho[:, :-2]) / 2.0
        return (self._gradient_angle(gx, gy, top, cols[0], +1),
                self._gradient_angle(gx, gy, top, cols[-1], -1))

    @staticmethod
    def _gradient_angle(gx, gy, top, i0, step):
        """arccos(|g_y| / |g|) of the gradient averaged over up to three
        columns stepping into the drop from the contact column, on the first
        fluid layer; columns with |g| <= 1e-10 are skipped."""
        nx, ny = gx.shape
        j = int(top[i0]) + 1
        if j >= ny:
            return float("nan")
        sx = sy = 0.0
        count = 0
        for k in range(3):
            i = i0 + step * k
            if i < 0 or i >= nx:
                continue
            a, b = gx[i, j], gy[i, j]
            if np.sqrt(a ** 2 + b ** 2) > 1e-10:
                sx += a
                sy += b
                count += 1
        if count == 0:
            return float("nan")
        ax, ay = sx / count, sy / count
        mag = np.sqrt(ax ** 2 + ay ** 2)
        if mag < 1e-10:
            return float("nan")
        return float(np.degrees(np.arccos(np.clip(abs(ay) / mag, 0, 1))))

    # -- report ------------------------------------------------------------------
    def summary(self):
        h = self.history
        if not h["step"]:
            print("No data recorded.")
            return
        print(f"\nMonitor Summary ({len(h['step'])} records)")
        print(f"  rho: [{min(h['rho_min']):.4f}, {max(h['rho_max']):.4f}]")
        print(f"  |u|_max: {max(h['u_max']):.4e}")
        if self.initial_mass and self.initial_mass > 0:
            change = (h["mass"][-1] - self.initial_mass) / self.initial_mass
            print(f"  Mass conservation: {change:+.4e} (relative change)")
        films = [f for f in h["film_thickness"] if f != float("inf")]
        if films:
            print(f"  Min film thickness: {min(films):.6e}")
        left = [a for a in h["contact_angle_left"] if not np.isnan(a)]
        right = [a for a in h["contact_angle_right"] if not np.isnan(a)]
        if left:
            print(f"  Contact angle (L): {left[-1]:.1f}°")
        if right:
            print(f"  Contact angle (R): {right[-1]:.1f}°")
