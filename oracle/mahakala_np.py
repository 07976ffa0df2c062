"""NumPy-FP64 restatement of Mahakala's per-ray hot path (the parity oracle).

TEST INFRASTRUCTURE ONLY -- imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``; never by ``mahakala_amd``.

Every function follows the reference function it names (file:line in
``mahakala/``), keeping its order of operations and its quirks:

* the fixed-N scan of ``geodesic_integrator`` with its freeze rule and the
  ``argmax(all(dt == 0)) + 2`` trim (``geodesics.py:233-281``); the scan is
  stopped once every ray is frozen, which is exact because a frozen ray emits
  the same (state, 0) row on every later step;
* ``jacfwd(metric)`` is evaluated by forward-mode dual numbers over the same
  expression graph as ``metric`` (``geodesics.py:88-104,305``) and ``imetric``
  by an LU inverse (``numpy.linalg.inv``, as ``jnp.linalg.inv`` at ``:347``);
* ``heaviside`` in ``_quadratic`` (``:58-66``), the no-op ``.at`` calls in
  ``select_photons_integrator`` (``:374,376``) with the ``-1`` index wrapping
  to the last row, the out-of-domain sampler rule and JAX's clamping of
  out-of-bounds gathers (``athenak.py:737-757``), the backward explicit-Euler
  transfer scan (``transfer.py:106-119``);
* the AthenaK ghost fill (``athenak.py:105-525``) including the stale-slice
  behaviour of its fine-to-coarse edge branch (``athenak.py:396-416``).

Parity status of each function (what pins it) is listed in the test files and
the design notes; nothing here is checked against a JAX run, which is not
possible in this environment (JAX is absent).
"""
from __future__ import annotations

import numpy as np

# --------------------------------------------------------------------------
# constants  (mahakala/constants.py:23-31, electrons.py:23-29, grmhd.py:26-32)
# --------------------------------------------------------------------------
EE = 4.8032e-10
KB = 1.3807e-16
CL = 2.99792458e10
ME = 9.1094e-28
MP = 1.6726e-24
HPL = 6.6261e-27
GNEWT = 6.6743e-8
Msun = 1.989e33


# --------------------------------------------------------------------------
# dual numbers: forward-mode derivative with 4 tangent directions (x^0..x^3)
# --------------------------------------------------------------------------
class _Dual:
    """Value ``v`` (shape S) with tangent ``d`` (shape S + (4,))."""

    __slots__ = ("v", "d")

    def __init__(self, v, d):
        self.v = v
        self.d = d

    @staticmethod
    def _lift(o, like):
        if isinstance(o, _Dual):
            return o
        o = np.asarray(o, dtype=np.float64)
        return _Dual(np.broadcast_to(o, like.v.shape), np.zeros(like.d.shape))

    def __add__(self, o):
        o = self._lift(o, self)
        return _Dual(self.v + o.v, self.d + o.d)

    __radd__ = __add__

    def __sub__(self, o):
        o = self._lift(o, self)
        return _Dual(self.v - o.v, self.d - o.d)

    def __rsub__(self, o):
        o = self._lift(o, self)
        return _Dual(o.v - self.v, o.d - self.d)

    def __neg__(self):
        return _Dual(-self.v, -self.d)

    def __mul__(self, o):
        o = self._lift(o, self)
        return _Dual(self.v * o.v, self.d * o.v[..., None] + self.v[..., None] * o.d)

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = self._lift(o, self)
        q = self.v / o.v
        return _Dual(q, (self.d - q[..., None] * o.d) / o.v[..., None])

    def __rtruediv__(self, o):
        o = self._lift(o, self)
        return o.__truediv__(self)

    def sqrt(self):
        s = np.sqrt(self.v)
        return _Dual(s, self.d / (2.0 * s)[..., None])


def _sqrt(x):
    return x.sqrt() if isinstance(x, _Dual) else np.sqrt(x)


# --------------------------------------------------------------------------
# metric  (geodesics.py:88-104)
# --------------------------------------------------------------------------
def _ks_f_l(x1, x2, x3, a):
    """f and covariant l_mu of g = eta + f l l, as the reference builds them."""
    aa = a * a
    zz = x3 * x3
    kk = 0.5 * (x1 * x1 + x2 * x2 + zz - aa)
    rr = _sqrt(kk * kk + aa * zz) + kk
    r = _sqrt(rr)
    f = (2.0 * rr * r) / (rr * rr + aa * zz)
    l1 = (r * x1 + a * x2) / (rr + aa)
    l2 = (r * x2 - a * x1) / (rr + aa)
    l3 = x3 / r
    return f, l1, l2, l3


def metric(x, bhspin):
    """g_{mu nu} at points x (..., 4) -> (..., 4, 4)   [geodesics.py:88-104]."""
    x = np.asarray(x, dtype=np.float64)
    f, l1, l2, l3 = _ks_f_l(x[..., 1], x[..., 2], x[..., 3], bhspin)
    one = np.ones_like(f)
    l = np.stack([one, l1, l2, l3], axis=-1)
    eta = np.diag([-1.0, 1.0, 1.0, 1.0])
    return eta + f[..., None, None] * (l[..., :, None] * l[..., None, :])


def jac_metric(x, bhspin):
    """d g_{mu nu} / d x^sigma, axis order [..., mu, nu, sigma] like jacfwd."""
    x = np.asarray(x, dtype=np.float64)
    shp = x.shape[:-1]
    comps = []
    for c in range(4):
        d = np.zeros(shp + (4,))
        d[..., c] = 1.0
        comps.append(_Dual(x[..., c].copy(), d))
    f, l1, l2, l3 = _ks_f_l(comps[1], comps[2], comps[3], bhspin)
    one = _Dual(np.ones(shp), np.zeros(shp + (4,)))
    l = [one, l1, l2, l3]
    jg = np.zeros(shp + (4, 4, 4))
    for m in range(4):
        for n in range(4):
            gmn = f * (l[m] * l[n])
            jg[..., m, n, :] = gmn.d
    return jg


def imetric(x, bhspin):
    """LU inverse of the metric   [geodesics.py:339-347]."""
    return np.linalg.inv(metric(x, bhspin))


# --------------------------------------------------------------------------
# camera  (geodesics.py:29-55, 107-134, 137-216, 219-230)
# --------------------------------------------------------------------------
def _Image_to_BH(x, y, z, i, d):
    i = i * np.pi / 180
    x_BH = -y * np.cos(i) + z * np.sin(i) + d * np.sin(i)
    y_BH = x
    z_BH = y * np.sin(i) + z * np.cos(i) + d * np.cos(i)
    return np.array([x_BH, y_BH, z_BH])


def _perpendicular(a):
    b = np.zeros_like(np.asarray(a, dtype=np.float64))
    b[0] = a[0] + a[1]
    b[1] = a[1] - a[0]
    return b


def _camera_from_plane(x, y, inclination, distance):
    n = x.size
    z = np.zeros(n)
    origin_BH = _Image_to_BH(0, 0, 0, inclination, distance)
    perp = _perpendicular([x, y])
    init_BH = _Image_to_BH(x, y, z, inclination, distance)
    perp_BH = _Image_to_BH(perp[0], perp[1], np.zeros(n), inclination, distance)
    vec1 = -init_BH.T + origin_BH
    vec2 = perp_BH - init_BH
    k_vec = np.cross(vec1, vec2.T)
    s0_x = np.array([np.zeros(n), init_BH[0].ravel(), init_BH[1].ravel(), init_BH[2].ravel()])
    s0_v = np.array([np.ones(n), k_vec.T[0].ravel(), k_vec.T[1].ravel(), k_vec.T[2].ravel()])
    return s0_x, s0_v


def get_initial_grid(inclination, distance, fov_lower, fov_upper, spacing, camera_type):
    kind = camera_type.lower()
    if kind == "grid":
        grid = np.linspace(fov_lower, fov_upper, 2 * spacing + 1)[1::2]
        x, y = np.meshgrid(grid, grid, indexing="ij")
        return _camera_from_plane(x.ravel(), y.ravel(), inclination, distance)
    if kind == "equator":
        grid = np.linspace(fov_lower, fov_upper, 2 * spacing + 1)[1::2]
        s0_x = np.zeros((4, len(grid)))
        s0_x[1] = distance
        s0_x[2] = grid
        s0_v = np.ones((4, len(grid)))
        s0_v[2] = 0
        s0_v[3] = 0
        return s0_x, s0_v
    return None


def get_camera_pixel(inclination, distance, radius, angle):
    size = np.size(radius)
    angle = np.asarray(angle, dtype=np.float64)
    radius = np.asarray(radius, dtype=np.float64)
    x = np.ones(size) * np.cos(angle) * radius
    y = np.ones(size) * np.sin(angle) * radius
    return _camera_from_plane(x, y, inclination, distance)


def _quadratic(A, b, C):
    bb = b * b
    AC = A * C
    dd = np.where(~np.isclose(bb, AC), bb - AC, 0.0)
    bs = np.heaviside(b, 1)
    with np.errstate(invalid="ignore", divide="ignore"):
        D = -(b + bs * np.sqrt(dd))
        x1 = D / A
        x2 = C / D
    return np.minimum(x1, x2), np.maximum(x1, x2)


def initial_condition(s0_x, s0_v, bhspin):
    """Null-normalise camera wavevectors, k^t fixed   [geodesics.py:69-85,219-230]."""
    xs = np.asarray(s0_x, dtype=np.float64).T  # (n, 4)
    vs = np.asarray(s0_v, dtype=np.float64).T
    g = metric(xs, bhspin)
    A = vs[:, 0] * g[:, 0, 0] * vs[:, 0]
    b = np.einsum("ni,ni->n", vs[:, 1:], g[:, 1:, 0]) * vs[:, 0]
    C = np.einsum("ni,nij,nj->n", vs[:, 1:], g[:, 1:, 1:], vs[:, 1:])
    d1, d2 = _quadratic(A, b, C)
    S = np.where(d1 > 0, d1, np.where(d2 > 0, d2, np.nan))
    out = np.empty((xs.shape[0], 8))
    out[:, :4] = xs
    out[:, 4] = vs[:, 0]
    with np.errstate(invalid="ignore", divide="ignore"):
        out[:, 5:] = vs[:, 1:] / S[:, None]
    return out


def initialize_geodesics_at_camera(bhspin, inclination, distance, fov_lower, fov_upper,
                                   pixels_per_side, camera_type="grid"):
    s0_x, s0_v = get_initial_grid(inclination, distance, fov_lower, fov_upper,
                                  pixels_per_side, camera_type)
    return initial_condition(s0_x, s0_v, bhspin)


# --------------------------------------------------------------------------
# geodesic equation and integrator  (geodesics.py:233-351)
# --------------------------------------------------------------------------
def radius_cal(x, bhspin):
    x = np.asarray(x, dtype=np.float64)
    R = np.sqrt(x[..., 1] ** 2 + x[..., 2] ** 2 + x[..., 3] ** 2)
    return np.sqrt((R ** 2 - bhspin ** 2
                    + np.sqrt((R ** 2 - bhspin ** 2) ** 2 + 4 * bhspin ** 2 * x[..., 3] ** 2)) / 2)


def radius_EH(a_spin):
    return 1 + np.sqrt(1 - a_spin ** 2)


def rhs(state, bhspin):
    """[v, -Gamma v v] for states (..., 8)   [geodesics.py:294-314]."""
    state = np.asarray(state, dtype=np.float64)
    x = state[..., :4]
    v = state[..., 4:]
    ig = imetric(x, bhspin)
    jg = jac_metric(x, bhspin)
    t1 = np.einsum("...mns,...s->...mn", jg, v)
    t1 = np.einsum("...mn,...n->...m", t1, v)
    t2 = np.einsum("...n,...mns->...ms", v, jg)
    t2 = np.einsum("...m,...ms->...s", v, t2)
    acc = np.einsum("...ab,...b->...a", ig, -t1 + 0.5 * t2)
    return np.concatenate([v, acc], axis=-1)


def RK4_gen(state1, dt, bhspin):
    state1 = np.asarray(state1, dtype=np.float64)
    dt = np.asarray(dt, dtype=np.float64).reshape(-1, 1)
    with np.errstate(invalid="ignore", over="ignore"):
        k1 = dt * rhs(state1, bhspin)
        k2 = dt * rhs(state1 + 0.5 * k1, bhspin)
        k3 = dt * rhs(state1 + 0.5 * k2, bhspin)
        k4 = dt * rhs(state1 + k3, bhspin)
        return state1 + 1 / 6 * (k1 + 2 * k2 + 2 * k3 + k4)


def _step_rule(state, bhspin, div, tol):
    with np.errstate(invalid="ignore"):
        dt = -(radius_cal(state[:, :4], bhspin) - radius_EH(bhspin)) / div
        bad = np.isnan(dt) | (np.abs(dt) * div < tol) | (np.abs(dt) * div > 1500)
    return np.where(bad, 0.0, dt)


def geodesic_integrator(N, s0, div, tol, bhspin, return_steps=False):
    """Fixed-N RK4 scan with freeze rule and trim   [geodesics.py:233-281].

    Returns (S (T, npx, 8), final_dt (T, npx)); with ``return_steps`` also the
    per-ray first frozen row K (N when the ray never froze).
    """
    s = np.array(s0, dtype=np.float64)
    npx = s.shape[0]
    rows_s, rows_dt = [], []
    first_zero = None
    K = np.full(npx, N, dtype=np.int64)
    for n in range(N):
        dt = _step_rule(s, bhspin, div, tol)
        new = RK4_gen(s, dt, bhspin)
        dtnew = _step_rule(new, bhspin, div, tol)
        rej = dtnew == 0.0
        dt = np.where(rej, 0.0, dt)
        new = np.where(rej[:, None], s, new)
        rows_s.append(s)
        rows_dt.append(dt)
        newly = (dt == 0.0) & (K == N)
        K[newly] = n
        if first_zero is None and np.all(dt == 0.0):
            first_zero = n
        s = new
        if first_zero is not None and (first_zero == 0 or n >= first_zero + 1):
            break
    if first_zero is None or first_zero < 1:
        T = N
    else:
        T = min(first_zero + 2, N)
    S = np.stack(rows_s[:T]) if len(rows_s) >= T else None
    D = np.stack(rows_dt[:T]) if len(rows_dt) >= T else None
    if S is None:  # all rays frozen from row 0: every row is (s0, 0)
        S = np.broadcast_to(rows_s[0], (T,) + rows_s[0].shape).copy()
        D = np.zeros((T, npx))
    if return_steps:
        return S, D, K
    return S, D


# --------------------------------------------------------------------------
# shadow bisection  (geodesics.py:354-435)
# --------------------------------------------------------------------------
def select_photons_integrator(inc, angle, radius, bhspin, distance=1000, max_steps=2000):
    s0_x, s0_v = get_camera_pixel(inc, distance, radius, angle)
    init_one = initial_condition(s0_x, s0_v, bhspin)
    S, dt = geodesic_integrator(max_steps, init_one, 40, 1e-2, bhspin)
    r = radius_cal(S, bhspin)
    maxi = np.argmax(dt, axis=0)
    maxi = maxi - 1  # the two .at[...] lines of the reference discard their result
    return r[maxi, np.arange(r.shape[1])]


def find_shadow_bisection_angles(bhspin, inc, angles, max_steps=2000, error_allowed=0.001,
                                 max_it=40):
    angles = np.asarray(angles, dtype=np.float64)
    inner = np.zeros_like(angles) + 0.5
    outer = np.zeros_like(angles) + 10
    error = outer - inner
    counter = 0
    while np.max(error) > error_allowed and counter < max_it:
        final = select_photons_integrator(inc, angles, (outer - inner) / 2 + inner, bhspin,
                                          max_steps=max_steps)
        fell = np.where(final < 100)
        away = np.where(final >= 100)
        inner[fell] = (outer[fell] - inner[fell]) / 2 + inner[fell]
        outer[away] = (outer[away] - inner[away]) / 2 + inner[away]
        error = outer - inner
        counter += 1
    return inner


def find_shadow_bisection(bhspin, inc, num_angles, max_steps=2000, error_allowed=0.001, max_it=40):
    angles = np.arange(num_angles) / num_angles * 2. * np.pi
    radii = find_shadow_bisection_angles(bhspin, inc, angles, max_it=max_it,
                                         error_allowed=error_allowed, max_steps=max_steps)
    return np.append(angles, angles[0]), np.append(radii, radii[0])


# --------------------------------------------------------------------------
# electrons, units, transfer  (electrons.py:32-50, grmhd.py:40-55, transfer.py:30-144)
# --------------------------------------------------------------------------
def rlow_rhigh_model(dens, u, beta, r_low=1, r_high=40, electron_gamma=4. / 3, ion_gamma=5. / 3):
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        T_ratio = (r_high * beta ** 2 + r_low) / (1 + beta ** 2)
        t_e = CL ** 2 * (MP * u * (electron_gamma - 1.) * (ion_gamma - 1.))
        t_e = t_e / (dens * ((ion_gamma - 1.) + (electron_gamma - 1.) * T_ratio))
        return t_e / (ME * CL * CL)


def get_units(M_BH, mass_scale):
    L_unit = GNEWT * M_BH / CL ** 2
    T_unit = L_unit / CL
    dens_unit = mass_scale / L_unit ** 3
    Ne_unit = dens_unit / (MP + ME)
    B_unit = CL * np.sqrt(4. * np.pi * dens_unit)
    return dict(L_unit=L_unit, T_unit=T_unit, dens_unit=dens_unit, Ne_unit=Ne_unit, B_unit=B_unit)


def synchrotron_coefficients(Ne, Theta_e, B, pitch_angle, nu, invariant=True, rescale_nu=1.):
    Ne, Theta_e, B, pitch_angle, nu = np.broadcast_arrays(
        *[np.asarray(v, dtype=np.float64) for v in (Ne, Theta_e, B, pitch_angle, nu)])
    with np.errstate(invalid="ignore", divide="ignore", over="ignore", under="ignore"):
        nuc = EE * B / (2. * np.pi * ME * CL)
        nus = (2. / 9.) * nuc * Theta_e ** 2 * np.sin(pitch_angle)
        X = nu / nus
        var = np.exp(-X ** (1. / 3))
        term = np.sqrt(X) + 2.0 ** (11. / 12) * X ** (1. / 6)
        em = Ne * nus * term ** 2 / (2. * Theta_e ** 2.)
        em = em * var * np.sqrt(2) * np.pi * EE ** 2 / (3.0 * CL)
        em = np.where(X > 1.e12, 0.0, em)
        em = np.where(Theta_e < 0.3, 0.0, em)
        bx = HPL * nu / (ME * CL * CL * Theta_e)
        series = bx / 24. * (24. + bx * (12. + bx * (4. + bx)))
        B_den = np.where(bx < 2.e-3, series, np.exp(bx) - 1)
        B_nu = (2. * HPL * nu ** 3. / B_den) / CL ** 2.
        ab = em / B_nu
        if invariant:
            rnu = nu * rescale_nu
            em = em / rnu ** 2.
            ab = ab * rnu
        em = np.where(np.isnan(em), 0.0, em)
        ab = np.where(np.isnan(ab), 0.0, ab)
    return em, ab


def solve_specific_intensity(emissivity, absorptivity, dt, L_unit, dIs=False):
    j = np.asarray(emissivity, dtype=np.float64)
    al = np.asarray(absorptivity, dtype=np.float64)
    dt = np.asarray(dt, dtype=np.float64)
    nsteps, npx = j.shape
    I = np.zeros(npx)
    out = []
    with np.errstate(invalid="ignore", over="ignore"):
        for i in range(nsteps - 1, 0, -1):
            dI = -dt[i - 1, :] * L_unit * (j[i, :] - (al[i] * I))
            I = I + dI
            out.append(dI)
    if dIs:
        return I, (np.stack(out) if out else np.zeros((0, npx)))
    return I


def solve_attenuated_emissivity(emissivity, absorptivity, dt, L_unit):
    j = np.asarray(emissivity, dtype=np.float64)
    al = np.asarray(absorptivity, dtype=np.float64)
    dt = np.asarray(dt, dtype=np.float64)
    nsteps, npx = j.shape
    tau = np.zeros(npx)
    out = []
    with np.errstate(invalid="ignore", over="ignore"):
        for i in range(1, nsteps):
            src = -j[i, :] * dt[i - 1, :] * L_unit
            dtau = al[i] * dt[i - 1] * L_unit
            out.append(np.exp(-tau) * src)
            tau = tau - dtau
    return np.stack(out) if out else np.zeros((0, npx))


# --------------------------------------------------------------------------
# AthenaK meshblock layout with ghost fill  (athenak.py:67-525)
# --------------------------------------------------------------------------
def _coarser_key(level, ti, tj, tk, di, dj, dk):
    return (level - 1, (ti + di) // 2, (tj + dj) // 2, (tk + dk) // 2)


def _finer_key(level, ti, tj, tk, di, dj, dk):
    return (level + 1, 2 * (ti + di), 2 * (tj + dj), 2 * (tk + dk))


def _halo_index(d, n):
    """Target index along one axis of the padded block for direction d."""
    if d == 1:
        return [n + 1]
    if d == -1:
        return [0]
    return list(range(1, n + 1))


def _fill_same(blk, src, di, dj, dk, n):
    """Copy from the same-level neighbour (athenak.py:212-229)."""
    def s(d):
        return [0] if d == 1 else ([n - 1] if d == -1 else list(range(n)))
    T = np.ix_(_halo_index(dk, n), _halo_index(dj, n), _halo_index(di, n))
    S = np.ix_(s(dk), s(dj), s(di))
    blk[:, T[0], T[1], T[2]] = src[:, S[0], S[1], S[2]]


def _fill_coarser(blk, src, t, d, n):
    """Injection from the coarser neighbour (athenak.py:231-292)."""
    idx = []
    for ax in range(3):
        odd = (t[ax] + d[ax]) % 2
        if d[ax] == 1:
            idx.append(([n + 1], [n // 2 if odd else 0]))
        elif d[ax] == -1:
            idx.append(([0], [n - 1 if odd else n // 2 - 1]))
        else:
            tgt = list(range(1, n + 1))
            base = n // 2 if odd else 0
            idx.append((tgt, [base + (p - 1) // 2 for p in tgt]))
    (tgi, sgi), (tgj, sgj), (tgk, sgk) = idx
    T = np.ix_(tgk, tgj, tgi)
    Sx = np.ix_(sgk, sgj, sgi)
    blk[:, T[0], T[1], T[2]] = src[:, Sx[0], Sx[1], Sx[2]]


def _fill_finer(blk, finer, key, t, d, n):
    """2x2x2 restriction from finer neighbours (athenak.py:303-514).

    The face and corner branches average the 8 fine cells under each halo
    cell.  The edge branch (exactly one of di, dj, dk zero) reproduces the
    reference's loop, where the sliced axis reads the selector left by the
    previous (v1, v2, v3) iteration: the first iteration reads fine cell 0
    for every halo cell, later ones the even or odd cells of the previous
    iteration's parity.
    """
    level, ni, nj, nk = key
    nsl = sum(1 for x in d if x == 0)
    base = [ni, nj, nk]
    src0 = [0, 0, 0]
    for ax in range(3):
        if d[ax] == -1:
            base[ax] += 1
            src0[ax] = n - 2
    h = n // 2

    def add(tgt_axes, child_off, sel_fn):
        # tgt_axes: per axis list of target indices; sel_fn(it, v) -> per axis source lists
        ck = (level, base[0] + child_off[0], base[1] + child_off[1], base[2] + child_off[2])
        src = finer[ck]
        it = 0
        for v1 in range(2):
            for v2 in range(2):
                for v3 in range(2):
                    sel = sel_fn(it, (v1, v2, v3))
                    T = np.ix_(tgt_axes[2], tgt_axes[1], tgt_axes[0])
                    S = np.ix_(sel[2], sel[1], sel[0])
                    vals = src[:, S[0], S[1], S[2]]
                    blk[:, T[0], T[1], T[2]] += np.broadcast_to(
                        (vals / 8.).astype(np.float64), blk[:, T[0], T[1], T[2]].shape)
                    it += 1

    if nsl == 0:
        tgt = [[n + 1] if d[ax] == 1 else [0] for ax in range(3)]
        add(tgt, (0, 0, 0), lambda it, v: [[src0[ax] + v[ax]] for ax in range(3)])
    elif nsl == 1:
        A = [ax for ax in range(3) if d[ax] == 0][0]
        for e in range(2):
            tgt = []
            off = [0, 0, 0]
            for ax in range(3):
                if d[ax] != 0:
                    tgt.append([n + 1] if d[ax] == 1 else [0])
                else:
                    tgt.append(list(range(1 + e * h, 1 + (e + 1) * h)))
                    off[ax] = e

            def sel(it, v, A=A):
                out = []
                for ax in range(3):
                    if ax != A:
                        out.append([src0[ax] + v[ax]])
                    elif it == 0:
                        out.append([0])  # the initial slice(1) selector: cell 0 only
                    else:
                        prev = it - 1
                        pv = ((prev >> 2) & 1, (prev >> 1) & 1, prev & 1)[A]
                        out.append(list(range(pv, n + pv, 2)))
                return out
            add(tgt, tuple(off), sel)
    else:
        sl = [ax for ax in range(3) if d[ax] == 0]
        for e1 in range(2):
            for e2 in range(2):
                e = {sl[0]: e1, sl[1]: e2}
                tgt = []
                off = [0, 0, 0]
                for ax in range(3):
                    if d[ax] != 0:
                        tgt.append([n + 1] if d[ax] == 1 else [0])
                    else:
                        tgt.append(list(range(1 + e[ax] * h, 1 + (e[ax] + 1) * h)))
                        off[ax] = e[ax]
                add(tgt, tuple(off),
                    lambda it, v: [[src0[ax] + v[ax]] if d[ax] != 0
                                   else list(range(v[ax], n + v[ax], 2)) for ax in range(3)])


def build_meshblocks(uov, B, levels, logical):
    """Ghost-padded (nmb, 8, n+2, n+2, n+2) array   [athenak.py:105-158].

    ``uov`` is (5, nmb, n, n, n) and ``B`` (3, nmb, n, n, n), as the AthenaK
    datasets; blocks are emitted in insertion order of the (level, i, j, k)
    map, which is the file order for a file without duplicate keys.
    """
    uov = np.asarray(uov)
    B = np.asarray(B)
    nprim, nmb, nk, nj, ni = uov.shape
    if not (nk == nj == ni):
        raise ValueError("cubic meshblocks only")
    n = ni
    index = {}
    for mb in range(nmb):
        index[(int(levels[mb]),) + tuple(int(v) for v in logical[mb])] = mb
    fields = np.concatenate([uov, B], axis=0)  # (8, nmb, n, n, n), file dtype
    out = np.zeros((nmb, 8, n + 2, n + 2, n + 2))
    for key, mb in index.items():
        level, ti, tj, tk = key
        blk = out[mb]
        blk[:, 1:-1, 1:-1, 1:-1] = fields[:, mb]
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for dk in (-1, 0, 1):
                    if di == 0 and dj == 0 and dk == 0:
                        continue
                    same = (level, ti + di, tj + dj, tk + dk)
                    if same in index:
                        _fill_same(blk, fields[:, index[same]], di, dj, dk, n)
                        continue
                    ck = _coarser_key(level, ti, tj, tk, di, dj, dk)
                    if ck in index:
                        _fill_coarser(blk, fields[:, index[ck]], (ti, tj, tk), (di, dj, dk), n)
                        continue
                    fk = _finer_key(level, ti, tj, tk, di, dj, dk)
                    if fk in index:
                        finer = {k: fields[:, v] for k, v in index.items() if k[0] == level + 1}
                        _fill_finer(blk, finer, fk, (ti, tj, tk), (di, dj, dk), n)
    return out


# --------------------------------------------------------------------------
# fluid sampler  (athenak.py:639-812)
# --------------------------------------------------------------------------
class Snapshot:
    """In-memory stand-in for AthenakFluidModel's loaded state."""

    def __init__(self, all_meshblocks, x1v, x2v, x3v, x1f, x2f, x3f, variable_names, bhspin,
                 fluid_gamma=None):
        self.all_meshblocks = np.asarray(all_meshblocks, dtype=np.float64)
        self.x1v, self.x2v, self.x3v = (np.asarray(v, dtype=np.float64) for v in (x1v, x2v, x3v))
        self.x1f, self.x2f, self.x3f = (np.asarray(v, dtype=np.float64) for v in (x1f, x2f, x3f))
        self.variable_names = [v.lower().strip() for v in variable_names]
        self.bhspin = bhspin
        self.fluid_gamma = fluid_gamma

    def idx(self, name):
        name = name.lower().strip()
        return self.variable_names.index(name) if name in self.variable_names else -1


def block_index(snap, pos):
    """Brute-force meshblock id per point, last match wins, -1 outside."""
    nmb = snap.all_meshblocks.shape[0]
    mb = np.full(pos.shape[:-1], -1, dtype=np.int64)
    for b in range(nmb):
        m = (snap.x1f[b][0] < pos[..., 1]) & (pos[..., 1] <= snap.x1f[b][-1])
        m &= (snap.x2f[b][0] < pos[..., 2]) & (pos[..., 2] <= snap.x2f[b][-1])
        m &= (snap.x3f[b][0] < pos[..., 3]) & (pos[..., 3] <= snap.x3f[b][-1])
        mb[m] = b
    return mb


def sample_prims(snap, S):
    """Trilinear primitives (..., 8) in file variable order; zero outside."""
    S = np.asarray(S, dtype=np.float64)
    mb = block_index(snap, S)
    n2 = snap.all_meshblocks.shape[-1]
    x_left = [snap.x1v[:, 0][mb], snap.x2v[:, 0][mb], snap.x3v[:, 0][mb]]
    dx = [(snap.x1v[:, 1] - snap.x1v[:, 0])[mb], (snap.x2v[:, 1] - snap.x2v[:, 0])[mb],
          (snap.x3v[:, 1] - snap.x3v[:, 0])[mb]]
    ids, fr = [], []
    with np.errstate(invalid="ignore"):
        for ax in range(3):
            q = S[..., ax + 1] - x_left[ax] + dx[ax]
            fr.append(np.remainder(q / dx[ax], 1.))
            i = np.floor_divide(q, dx[ax])
            i = np.where(np.isfinite(i), i, 0).astype(np.int64)
            ids.append(i)
    data = snap.all_meshblocks
    i1, i2, i3 = ids

    def g(a3, a2, a1):
        c3 = np.clip(i3 + a3, 0, n2 - 1)
        c2 = np.clip(i2 + a2, 0, n2 - 1)
        c1 = np.clip(i1 + a1, 0, n2 - 1)
        return np.moveaxis(data[mb, :, c3, c2, c1], -1, -1)

    d1, d2, d3 = (f[..., None] for f in fr)
    with np.errstate(invalid="ignore"):
        daa = g(0, 0, 0) + (g(0, 0, 1) - g(0, 0, 0)) * d1
        dab = g(0, 1, 0) + (g(0, 1, 1) - g(0, 1, 0)) * d1
        dba = g(1, 0, 0) + (g(1, 0, 1) - g(1, 0, 0)) * d1
        dbb = g(1, 1, 0) + (g(1, 1, 1) - g(1, 1, 0)) * d1
        da = daa + (dab - daa) * d2
        db = dba + (dbb - dba) * d2
        prims = da + (db - da) * d3
    prims = np.where((mb == -1)[..., None], 0.0, prims)
    return prims, mb


def four_vectors(snap, S):
    """u^mu, u_mu, b^mu, b_mu and the metric at the points of S (athenak.py:759-781)."""
    S = np.asarray(S, dtype=np.float64)
    iU1, iB1 = snap.idx("velx"), snap.idx("bcc1")
    prims, _ = sample_prims(snap, S)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        gcov = metric(S[..., :4], snap.bhspin)
        gcon = np.linalg.inv(gcov)
        alpha = np.sqrt(1. / (-gcon[..., 0, 0]))
        U = prims[..., iU1:iU1 + 3]
        gamma = np.sqrt(1 + np.einsum("...j,...j->...", np.einsum("...i,...ij->...j", U, gcov[..., 1:, 1:]), U))
        ucon = np.empty(S.shape[:-1] + (4,))
        ucon[..., 0] = gamma / alpha
        for c in range(3):
            ucon[..., c + 1] = U[..., c] - gamma * alpha * gcon[..., 0, c + 1]
        ucov = np.einsum("...ij,...j->...i", gcov, ucon)
        Bp = prims[..., iB1:iB1 + 3]
        bcon = np.empty_like(ucon)
        bcon[..., 0] = np.einsum("...i,...i->...", Bp, ucov[..., 1:])
        for c in range(3):
            bcon[..., c + 1] = (Bp[..., c] + ucon[..., c + 1] * bcon[..., 0]) / ucon[..., 0]
        bcov = np.einsum("...ij,...j->...i", gcov, bcon)
    return ucon, ucov, bcon, bcov, gcov


def get_fluid_scalars_from_geodesics(snap, S, fallback_pitch_angle=np.pi / 3.):
    S = np.asarray(S, dtype=np.float64)
    iU1, iU2, iU3 = snap.idx("velx"), snap.idx("vely"), snap.idx("velz")
    iB1, iB2, iB3 = snap.idx("bcc1"), snap.idx("bcc2"), snap.idx("bcc3")
    if iU2 != iU1 + 1 or iU3 != iU1 + 2:
        raise ValueError("Velocity indices are not as expected")
    if iB2 != iB1 + 1 or iB3 != iB1 + 2:
        raise ValueError("Magnetic field indices are not as expected")
    prims, _ = sample_prims(snap, S)
    X = S[..., :4]
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        gcov = metric(X, snap.bhspin)
        gcon = np.linalg.inv(gcov)
        alpha = np.sqrt(1. / (-gcon[..., 0, 0]))
        U = prims[..., iU1:iU1 + 3]
        gamma = np.sqrt(1 + np.einsum("...j,...j->...",
                                      np.einsum("...i,...ij->...j", U, gcov[..., 1:, 1:]), U))
        ucon = np.empty(S.shape[:-1] + (4,))
        ucon[..., 0] = gamma / alpha
        for c in range(3):
            ucon[..., c + 1] = U[..., c] - gamma * alpha * gcon[..., 0, c + 1]
        ucov = np.einsum("...ij,...j->...i", gcov, ucon)
        Bp = prims[..., iB1:iB1 + 3]
        bcon = np.empty_like(ucon)
        bcon[..., 0] = np.einsum("...i,...i->...", Bp, ucov[..., 1:])
        for c in range(3):
            bcon[..., c + 1] = (Bp[..., c] + ucon[..., c + 1] * bcon[..., 0]) / ucon[..., 0]
        bcov = np.einsum("...ij,...j->...i", gcov, bcon)
        k = S[..., 4:]
        kdotu = np.einsum("...i,...i->...", k, ucov)
        kdotb = np.einsum("...i,...i->...", k, bcov)
        bdotb = np.einsum("...i,...i->...", bcon, bcov)
        p = kdotb / (np.abs(kdotu) * np.sqrt(bdotb))
        p = np.where(np.isnan(p), np.cos(fallback_pitch_angle), p)
        p = np.where(np.abs(p) > 1.0, p / np.abs(p), p)
        pitch = np.arccos(p)
        b = np.sqrt(bdotb)
    return dict(dens=prims[..., snap.idx("dens")], u=prims[..., snap.idx("eint")],
                pitch_angle=pitch, kdotu=kdotu, b=b)


# --------------------------------------------------------------------------
# image  (images.py:30-144)
# --------------------------------------------------------------------------
def image_from_history(snap, S, final_dt, fluid_gamma, mass_scale=1.e26, M_bh=6.2e9 * Msun,
                       r_high=40, observing_frequency=230.e9):
    """The per-chunk body of make_image (images.py:81-120) on one history."""
    fs = get_fluid_scalars_from_geodesics(snap, S)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        bsq = fs["b"] * fs["b"]
        beta = fs["u"] * (fluid_gamma - 1.) / bsq / 0.5
        sigma = bsq / fs["dens"]
        Theta_e = rlow_rhigh_model(fs["dens"], fs["u"], beta, r_high=r_high)
    units = get_units(M_bh, mass_scale)
    Ne = units["Ne_unit"] * fs["dens"]
    Bg = units["B_unit"] * fs["b"]
    local_nu = -fs["kdotu"] * observing_frequency
    em, ab = synchrotron_coefficients(Ne, Theta_e, Bg, fs["pitch_angle"], local_nu,
                                      invariant=True, rescale_nu=1. / observing_frequency)
    with np.errstate(invalid="ignore"):
        cut = sigma > 100.
    em = np.where(cut, 0.0, em)
    ab = np.where(cut, 0.0, ab)
    return solve_specific_intensity(em, ab, final_dt, units["L_unit"])


def make_image(snap, camera_inclination=60, camera_distance=1000, mass_scale=1.e26,
               M_bh=6.2e9 * Msun, r_high=40, observing_frequency=230.e9, fov=20, resolution=160,
               max_nsteps=10000):
    s0 = initialize_geodesics_at_camera(snap.bhspin, camera_inclination, camera_distance,
                                        -fov / 2., fov / 2., resolution)
    S, dt = geodesic_integrator(max_nsteps, s0, 40, 1e-4, snap.bhspin)
    I = image_from_history(snap, S, dt, snap.fluid_gamma, mass_scale, M_bh, r_high,
                           observing_frequency)
    return I.reshape((resolution, resolution))


# --------------------------------------------------------------------------
# Bardeen (1973) critical curve: closed form used to pin the shadow
# --------------------------------------------------------------------------
def bardeen_curve(a, inc_deg, r):
    """(angle, radius) on the image plane of the Kerr critical curve.

    xi = (r^2 (3 - r) - a^2 (r + 1)) / (a (r - 1)),
    eta = r^3 (4 a^2 - r (r - 3)^2) / (a^2 (r - 1)^2),
    alpha = -xi / sin i,  beta = +sqrt(eta + a^2 cos^2 i - xi^2 cot^2 i).
    """
    i = np.deg2rad(inc_deg)
    xi = (r ** 2 * (3 - r) - a ** 2 * (r + 1)) / (a * (r - 1))
    eta = r ** 3 * (4 * a ** 2 - r * (r - 3) ** 2) / (a ** 2 * (r - 1) ** 2)
    beta2 = eta + a ** 2 * np.cos(i) ** 2 - xi ** 2 / np.tan(i) ** 2
    ok = beta2 >= 0
    alpha = -xi[ok] / np.sin(i)
    beta = np.sqrt(beta2[ok])
    return np.arctan2(beta, alpha), np.hypot(alpha, beta)


def photon_orbit_radii(a):
    """Prograde and retrograde equatorial photon-orbit radii."""
    rp = 2 * (1 + np.cos(2. / 3 * np.arccos(-abs(a))))
    rm = 2 * (1 + np.cos(2. / 3 * np.arccos(abs(a))))
    return rp, rm
