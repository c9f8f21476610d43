"""Synthetic lattice fields in MILC's raw layout (numpy).

Raw layout (convention of the reference's create_*_from_raw, qphix_d3.cpp:79-174,
970-1085): all even sites first, then all odd sites; within a parity the
site index is i = xh + nxh*(y + ny*(z + nz*t)) with x = 2*xh + ((y+z+t+p)&1).
Colour vectors hold 6 reals per site ([colour][re,im]); link arrays hold 72
reals per site ([dir X,Y,Z,T][row][col][re,im]).

Links carry MILC's staggered phases eta_x=(-1)^t, eta_y=(-1)^(t+x),
eta_z=(-1)^(t+x+y), eta_t=1 and an antiperiodic time boundary (fat t-links at
t=nt-1, Naik t-links at t>=nt-3), folded in as the caller (MILC) does.  Back
links follow MILC's double-store convention: fatback_mu(x) = -U_mu(x-mu)^dagger,
lngback_mu(x) = -N_mu(x-3mu)^dagger, which makes D anti-Hermitian.
"""
import numpy as np


class Lattice:
    def __init__(self, dims):
        self.dims = tuple(int(d) for d in dims)
        nx, ny, nz, nt = self.dims
        assert all(d % 4 == 0 for d in self.dims), "extents must be multiples of 4"
        self.nxh = nx // 2
        self.V = nx * ny * nz * nt
        self.vh = self.V // 2
        i = np.arange(self.vh, dtype=np.int64)
        xh = i % self.nxh
        r = i // self.nxh
        y = r % ny
        r //= ny
        z = r % nz
        t = r // nz
        xs, ys, zs, ts = [], [], [], []
        for p in (0, 1):
            xs.append(2 * xh + ((y + z + t + p) & 1))
            ys.append(y)
            zs.append(z)
            ts.append(t)
        self.x = np.concatenate(xs)
        self.y = np.concatenate(ys)
        self.z = np.concatenate(zs)
        self.t = np.concatenate(ts)
        self.coords = np.stack([self.x, self.y, self.z, self.t])

    def index(self, x, y, z, t):
        nx, ny, nz, nt = self.dims
        x, y, z, t = x % nx, y % ny, z % nz, t % nt
        p = (x + y + z + t) & 1
        return p * self.vh + x // 2 + self.nxh * (y + ny * (z + nz * t))

    def shifted(self, mu, hop):
        """Raw index of site + hop * mu-hat for every raw site."""
        c = [self.x, self.y, self.z, self.t]
        c = [ci + (hop if d == mu else 0) for d, ci in enumerate(c)]
        return self.index(*c)

    def eta(self, mu):
        x, y, t = self.x, self.y, self.t
        if mu == 0:
            e = (-1) ** (t & 1)
        elif mu == 1:
            e = (-1) ** ((t + x) & 1)
        elif mu == 2:
            e = (-1) ** ((t + x + y) & 1)
        else:
            e = np.ones(self.V, dtype=np.int64)
        return e.astype(np.float64)

    def bc(self, mu, hop, antiperiodic=True):
        """-1 on links of direction mu and length hop that cross t = nt."""
        if mu != 3 or not antiperiodic:
            return np.ones(self.V)
        nt = self.dims[3]
        return np.where(self.t + hop >= nt, -1.0, 1.0)

    def parity_mask(self, p):
        m = np.zeros(self.V, dtype=bool)
        m[p * self.vh:(p + 1) * self.vh] = True
        return m


# ----------------------------------------------------------------- SU(3)
def dagger(u):
    return np.conj(np.swapaxes(u, -1, -2))


def haar_su3(rng, n):
    """Haar-random SU(3): QR of a complex Gaussian, diagonal phases fixed,
    determinant normalised."""
    z = (rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    q = q / (det ** (1.0 / 3.0))[:, None, None]
    return q


def weak_su3(rng, n, eps):
    """exp(i eps H), H Gaussian Hermitian traceless."""
    a = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
    h = 0.5 * (a + dagger(a))
    tr = np.trace(h, axis1=1, axis2=2).real / 3.0
    h = h - tr[:, None, None] * np.eye(3)[None]
    w, v = np.linalg.eigh(h)
    return (v * np.exp(1j * eps * w)[:, None, :]) @ dagger(v)


def near_identity(rng, n, sigma):
    """1 + sigma * complex Gaussian per entry (not unitary)."""
    g = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
    return np.eye(3)[None] + sigma * g


# ----------------------------------------------------------------- links
def fold_phases(lat, fwd, hop, antiperiodic=True):
    """Multiply the forward links (V,4,3,3) by eta_mu(x) and the t boundary sign."""
    out = fwd.copy()
    for mu in range(4):
        out[:, mu] *= (lat.eta(mu) * lat.bc(mu, hop, antiperiodic))[:, None, None]
    return out


def back_links(lat, fwd, hop):
    """MILC double-store back links: back_mu(x) = -fwd_mu(x - hop*mu)^dagger."""
    back = np.empty_like(fwd)
    for mu in range(4):
        back[:, mu] = -dagger(fwd[lat.shifted(mu, -hop), mu])
    return back


def to_raw(a, dtype=np.float64):
    """complex (..., 3, 3) or (..., 3) array -> flat raw real array."""
    return np.ascontiguousarray(np.ascontiguousarray(a).view(np.float64).astype(dtype, copy=False).ravel())


def from_raw_links(raw, lat):
    return raw.astype(np.float64).view(np.complex128).reshape(lat.V, 4, 3, 3)


def from_raw_vec(raw, lat):
    return raw.astype(np.float64).view(np.complex128).reshape(lat.V, 3)


def make_links(lat, kind, seed, eps=0.3, fat_sigma=None, antiperiodic=True):
    """Forward fat and long links (complex (V,4,3,3)), phases folded in.

    kind: "haar"  -- fat = 1 + 0.1 Gaussian, long = Haar SU(3) (8^4 config)
          "weak"  -- long = exp(i eps H), fat = 1 + 0.3 Gaussian (32^4, 32^3x64, 48^3x96)
    """
    rng = np.random.default_rng(seed)
    n = lat.V * 4
    if kind == "haar":
        fat = near_identity(rng, n, 0.1 if fat_sigma is None else fat_sigma)
        lng = haar_su3(rng, n)
    elif kind == "weak":
        fat = near_identity(rng, n, 0.3 if fat_sigma is None else fat_sigma)
        lng = weak_su3(rng, n, eps)
    else:
        raise ValueError(kind)
    fat = fold_phases(lat, fat.reshape(lat.V, 4, 3, 3), 1, antiperiodic)
    lng = fold_phases(lat, lng.reshape(lat.V, 4, 3, 3), 3, antiperiodic)
    return fat, lng


def free_links(lat, c1, c3, antiperiodic=True):
    """Free field: fat = c1 * eta * bc * 1, long = c3 * eta * bc * 1."""
    one = np.broadcast_to(np.eye(3, dtype=np.complex128), (lat.V, 4, 3, 3)).copy()
    fat = fold_phases(lat, c1 * one, 1, antiperiodic)
    lng = fold_phases(lat, c3 * one, 3, antiperiodic)
    return fat, lng


def raw_link_set(lat, fat, lng, dtype=np.float64):
    """The four raw arrays create_L_from_raw takes: fat, lng, fatback, lngback."""
    return (to_raw(fat, dtype), to_raw(lng, dtype),
            to_raw(back_links(lat, fat, 1), dtype), to_raw(back_links(lat, lng, 3), dtype))


def gaussian_source(lat, seed, parity=None):
    """Complex Gaussian colour vector (V,3); zero off `parity` (0 even, 1 odd) if given."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((lat.V, 3)) + 1j * rng.standard_normal((lat.V, 3))
    if parity is not None:
        v[~lat.parity_mask(parity)] = 0
    return v


def gauge_transform(lat, fat, lng, g):
    """U_mu(x) -> g(x) U_mu(x) g(x+hop mu)^dagger for fat (hop 1) and long (hop 3)."""
    f2 = np.empty_like(fat)
    l2 = np.empty_like(lng)
    for mu in range(4):
        f2[:, mu] = g @ fat[:, mu] @ dagger(g[lat.shifted(mu, 1)])
        l2[:, mu] = g @ lng[:, mu] @ dagger(g[lat.shifted(mu, 3)])
    return f2, l2


def plane_wave(lat, p, colour, parity):
    """exp(i p.x) in one colour on the sites of `parity`, zero elsewhere."""
    ph = np.exp(1j * (p[0] * lat.x + p[1] * lat.y + p[2] * lat.z + p[3] * lat.t))
    v = np.zeros((lat.V, 3), dtype=np.complex128)
    v[:, colour] = ph
    v[~lat.parity_mask(parity)] = 0
    return v


def gauge_field(lat, kind, seed, eps=0.3):
    """Plain SU(3) gauge links (no staggered phases), complex (V,4,3,3)."""
    rng = np.random.default_rng(seed)
    n = lat.V * 4
    u = haar_su3(rng, n) if kind == "haar" else weak_su3(rng, n, eps)
    return u.reshape(lat.V, 4, 3, 3)


def antihermitian_traceless(rng, shape):
    """Random traceless anti-Hermitian 3x3 matrices."""
    a = rng.standard_normal(shape + (3, 3)) + 1j * rng.standard_normal(shape + (3, 3))
    a = 0.5 * (a - dagger(a))
    tr = np.trace(a, axis1=-2, axis2=-1) / 3.0
    return a - tr[..., None, None] * np.eye(3)


def symanzik_coeffs(beta, u0):
    """Tadpole-improved one-loop Symanzik coefficients (Alford et al.,
    Phys. Lett. B361 (1995) 87), scaled by beta/3 so that
    S = -sum c Re Tr U_loop equals beta * sum c_i (1 - Re Tr U_loop / 3) up to a
    constant: (plaquette, rectangle, parallelogram)."""
    alpha_s = -4.0 * np.log(u0) / 3.06839
    c = (1.0, -1.0 / (20.0 * u0 * u0) * (1.0 + 0.4805 * alpha_s),
         -1.0 / (u0 * u0) * 0.03325 * alpha_s)
    return tuple(beta / 3.0 * ci for ci in c)


def free_eigenvalue(p, c1, c3, m2):
    """m2 + sum_mu (2 c1 sin p_mu + 2 c3 sin 3 p_mu)^2: eigenvalue of
    m2 - D_eo D_oe on an even-site plane wave for the free-field links."""
    return m2 + sum((2 * c1 * np.sin(q) + 2 * c3 * np.sin(3 * q)) ** 2 for q in p)
