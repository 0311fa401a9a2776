"""FP64 Ewald sums -- the checker of the PME kernels.

TEST INFRASTRUCTURE: only tests/ and bench.py's CPU leg use this file.

The reference declares a PME flag (gcmc/cpp/gcmc.h:91, gcmc/python/values.py:55)
that no kernel reads, so there is no reference algorithm to restate; this is
the textbook Ewald decomposition (Ewald 1921; notation of Essmann et al.,
J. Chem. Phys. 103, 8577 (1995)) of the Coulomb energy of point charges in an
orthorhombic periodic box, in the units of calcEnergy
(K = 1388.431112 kJ/mol A e^-2, gcmc/cpp/gcmc_energy.h:137):

    E      = E_real + E_recip + E_self
    E_real = K sum_{i<j, n} q_i q_j erfc(beta r) / r        (minimum image, r <= rc)
    E_recip= K / (2 pi V) sum_{m != 0} exp(-pi^2 m^2 / beta^2) / m^2 |S(m)|^2
    S(m)   = sum_j q_j exp(2 pi i m . x_j),  m = (k1/L1, k2/L2, k3/L3)
    E_self = -K beta / sqrt(pi) sum_j q_j^2

Pinned by: the NaCl Madelung constant, independence of the total from beta
(tests/test_pme_oracle.py).
"""

import math

import numpy as np
from scipy.special import erfc

KCOUL = 1388.431112


def recip_energy(xyz, q, L, beta, kmax):
    """Reciprocal-space Ewald energy with |k_i| <= kmax[i] (FP64)."""
    xyz = np.asarray(xyz, np.float64)
    q = np.asarray(q, np.float64)
    L = np.asarray(L, np.float64)
    V = float(np.prod(L))
    k = [np.arange(-kmax[i], kmax[i] + 1) for i in range(3)]
    K1, K2, K3 = np.meshgrid(*k, indexing='ij')
    m = np.stack([K1.ravel() / L[0], K2.ravel() / L[1], K3.ravel() / L[2]], 1)
    m2 = (m * m).sum(1)
    keep = m2 > 0
    m, m2 = m[keep], m2[keep]
    e = 0.0
    for c0 in range(0, len(m), 4096):
        mm = m[c0:c0 + 4096]
        ph = 2 * math.pi * xyz @ mm.T                       # [N, M]
        S_re = (q[:, None] * np.cos(ph)).sum(0)
        S_im = (q[:, None] * np.sin(ph)).sum(0)
        w = np.exp(-math.pi ** 2 * m2[c0:c0 + 4096] / beta ** 2) / m2[c0:c0 + 4096]
        e += float((w * (S_re ** 2 + S_im ** 2)).sum())
    return KCOUL * e / (2 * math.pi * V)


def real_energy(xyz, q, L, beta, rc):
    """Real-space Ewald energy, minimum image, pairs with r <= rc (FP64)."""
    xyz = np.asarray(xyz, np.float64)
    q = np.asarray(q, np.float64)
    L = np.asarray(L, np.float64)
    e = 0.0
    for i in range(len(q) - 1):
        d = xyz[i + 1:] - xyz[i]
        d -= np.round(d / L) * L
        r = np.sqrt((d * d).sum(1))
        m = r <= rc
        e += float((q[i] * q[i + 1:][m] * erfc(beta * r[m]) / r[m]).sum())
    return KCOUL * e


def self_energy(q, beta):
    q = np.asarray(q, np.float64)
    return -KCOUL * beta / math.sqrt(math.pi) * float((q * q).sum())


def ewald_energy(xyz, q, L, beta, rc, kmax):
    return real_energy(xyz, q, L, beta, rc) + recip_energy(xyz, q, L, beta, kmax) + self_energy(q, beta)


def recip_potential(xyz, q, L, beta, kmax, points):
    """Reciprocal-space potential (kJ/mol/e) of the charges at `points`:
    phi(r) = K / (pi V) sum_{m != 0} exp(-pi^2 m^2/beta^2)/m^2 Re(S(m) e^{-2 pi i m.r})."""
    xyz = np.asarray(xyz, np.float64)
    q = np.asarray(q, np.float64)
    pts = np.asarray(points, np.float64).reshape(-1, 3)
    L = np.asarray(L, np.float64)
    V = float(np.prod(L))
    k = [np.arange(-kmax[i], kmax[i] + 1) for i in range(3)]
    K1, K2, K3 = np.meshgrid(*k, indexing='ij')
    m = np.stack([K1.ravel() / L[0], K2.ravel() / L[1], K3.ravel() / L[2]], 1)
    m2 = (m * m).sum(1)
    keep = m2 > 0
    m, m2 = m[keep], m2[keep]
    out = np.zeros(len(pts))
    for c0 in range(0, len(m), 4096):
        mm = m[c0:c0 + 4096]
        w = np.exp(-math.pi ** 2 * m2[c0:c0 + 4096] / beta ** 2) / m2[c0:c0 + 4096]
        ph = 2 * math.pi * xyz @ mm.T
        S_re = (q[:, None] * np.cos(ph)).sum(0)
        S_im = (q[:, None] * np.sin(ph)).sum(0)
        pp = 2 * math.pi * pts @ mm.T
        out += (w[None, :] * (S_re[None, :] * np.cos(pp) + S_im[None, :] * np.sin(pp))).sum(1)
    return KCOUL * out / (math.pi * V)


def bspline4(u):
    """Cardinal B-spline of order 4, M4(u), u in [0, 4)."""
    u = np.asarray(u, np.float64)
    out = np.zeros_like(u)
    a = (u >= 0) & (u < 1)
    out[a] = u[a] ** 3 / 6
    b = (u >= 1) & (u < 2)
    ub = u[b]
    out[b] = (-3 * ub ** 3 + 12 * ub ** 2 - 12 * ub + 4) / 6
    c = (u >= 2) & (u < 3)
    uc = u[c]
    out[c] = (3 * uc ** 3 - 24 * uc ** 2 + 60 * uc - 44) / 6
    d = (u >= 3) & (u < 4)
    out[d] = (4 - u[d]) ** 3 / 6
    return out


def nacl_crystal(ncell=2, a=5.64):
    """Rock-salt lattice, ncell^3 conventional cells of edge a (A):
    returns (xyz, q, L)."""
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    pts, qs = [], []
    for i in range(ncell):
        for j in range(ncell):
            for k in range(ncell):
                o = np.array([i, j, k], np.float64)
                for b in base:
                    pts.append((o + b) * a)
                    qs.append(1.0)
                    pts.append((o + b + np.array([.5, 0, 0])) * a)
                    qs.append(-1.0)
    L = [ncell * a] * 3
    return np.array(pts), np.array(qs), L


class EwaldEnv:
    """Reciprocal Ewald potential of fixed charges, structure factors computed
    once: phi(points) = K/(pi V) sum_m w(m) Re(S(m) e^{-2 pi i m.r})."""

    def __init__(self, xyz, q, L, beta, kmax):
        xyz = np.asarray(xyz, np.float64)
        q = np.asarray(q, np.float64)
        L = np.asarray(L, np.float64)
        self.V = float(np.prod(L))
        k = [np.arange(-kmax[i], kmax[i] + 1) for i in range(3)]
        K1, K2, K3 = np.meshgrid(*k, indexing='ij')
        m = np.stack([K1.ravel() / L[0], K2.ravel() / L[1], K3.ravel() / L[2]], 1)
        m2 = (m * m).sum(1)
        keep = m2 > 0
        self.m = m[keep]
        m2 = m2[keep]
        self.w = np.exp(-math.pi ** 2 * m2 / beta ** 2) / m2
        self.S_re = np.zeros(len(self.m))
        self.S_im = np.zeros(len(self.m))
        for c0 in range(0, len(q), 256):
            ph = 2 * math.pi * xyz[c0:c0 + 256] @ self.m.T
            self.S_re += (q[c0:c0 + 256, None] * np.cos(ph)).sum(0)
            self.S_im += (q[c0:c0 + 256, None] * np.sin(ph)).sum(0)

    def phi(self, points):
        pts = np.asarray(points, np.float64).reshape(-1, 3)
        pp = 2 * math.pi * pts @ self.m.T
        return KCOUL / (math.pi * self.V) * (self.w[None, :] * (self.S_re[None, :] * np.cos(pp) +
                                                                self.S_im[None, :] * np.sin(pp))).sum(1)


def _spline_points(x, L, K):
    """Grid points and order-4 cardinal B-spline weights of positions x along
    one axis (scaled coordinate u = K frac(x / L), points floor(u) - i with
    weight M4(u - floor(u) + i), i = 0..3), FP64."""
    fr = np.asarray(x, np.float64) / L
    u = (fr - np.floor(fr)) * K
    u = np.where(u >= K, 0.0, u)
    f = np.floor(u)
    t = u - f
    idx = (f.astype(np.int64)[:, None] - np.arange(4)[None, :]) % K
    w = bspline4(t[:, None] + np.arange(4)[None, :])
    return idx, w


class PMEEnv:
    """Smooth particle-mesh Ewald (Essmann et al. 1995, order-4 B-splines) in
    FP64: the same discretisation as the GPU's PME kernels (charge grid Q,
    influence function G with exact B-spline moduli, potential grid
    Phi = IFFT(G FFT(Q)) unnormalised), so that a comparison with them
    measures arithmetic error alone; its discretisation error against the
    Ewald sums (EwaldEnv) is pinned separately."""

    def __init__(self, xyz, q, L, beta, K):
        xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
        q = np.asarray(q, np.float64)
        self.L = np.asarray(L, np.float64)
        self.K = tuple(int(k) for k in K)
        K0, K1, K2 = self.K
        Q = np.zeros(K0 * K1 * K2)
        i0, w0 = _spline_points(xyz[:, 0], self.L[0], K0)
        i1, w1 = _spline_points(xyz[:, 1], self.L[1], K1)
        i2, w2 = _spline_points(xyz[:, 2], self.L[2], K2)
        for a in range(4):
            for b in range(4):
                flat = (i0[:, a, None] * K1 + i1[:, b, None]) * K2 + i2
                Q += np.bincount(flat.ravel(), weights=(q[:, None] * w0[:, a, None] * w1[:, b, None] *
                                                        w2).ravel(), minlength=len(Q))
        Q = Q.reshape(self.K)
        V = float(np.prod(self.L))

        def freq(n, Ln, half=False):
            m = np.arange(n // 2 + 1 if half else n)
            if not half:
                m = np.where(m <= n // 2, m, m - n)
            return m / Ln

        def bmod2(n, count):
            m = np.arange(count)
            k = np.arange(3)
            M4 = np.array([1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0])
            ph = 2.0 * math.pi * m[:, None] * k[None, :] / n
            re = (M4 * np.cos(ph)).sum(1)
            im = (M4 * np.sin(ph)).sum(1)
            return 1.0 / (re * re + im * im)

        ma = freq(K0, self.L[0])[:, None, None]
        mb = freq(K1, self.L[1])[None, :, None]
        mc = freq(K2, self.L[2], half=True)[None, None, :]
        m2 = ma * ma + mb * mb + mc * mc
        B = bmod2(K0, K0)[:, None, None] * bmod2(K1, K1)[None, :, None] * bmod2(K2, K2 // 2 + 1)[None, None, :]
        with np.errstate(divide='ignore', invalid='ignore'):
            G = KCOUL / (math.pi * V) * np.exp(-math.pi ** 2 * m2 / beta ** 2) / m2 * B
        G[0, 0, 0] = 0.0
        X = np.fft.rfftn(Q)
        self.energy = 0.5 * float((G * np.abs(X) ** 2 * np.where(
            (np.arange(K2 // 2 + 1) == 0) | (2 * np.arange(K2 // 2 + 1) == K2), 1.0, 2.0)[None, None, :]).sum())
        self.Phi = np.fft.irfftn(G * X, s=self.K, axes=(0, 1, 2)) * (K0 * K1 * K2)

    def phi(self, points):
        pts = np.asarray(points, np.float64).reshape(-1, 3)
        K0, K1, K2 = self.K
        i0, w0 = _spline_points(pts[:, 0], self.L[0], K0)
        i1, w1 = _spline_points(pts[:, 1], self.L[1], K1)
        i2, w2 = _spline_points(pts[:, 2], self.L[2], K2)
        out = np.zeros(len(pts))
        for a in range(4):
            for b in range(4):
                out += w0[:, a] * w1[:, b] * (w2 * self.Phi[i0[:, a, None], i1[:, b, None], i2]).sum(1)
        return out


def present_groups(sysd, exclude=-1):
    """Residue indices in the state: fixed groups, then each species' active
    slots (calcEnergy's order, gcmc/cpp/gcmc_energy.h:152-191), minus exclude."""
    frag = sysd['frag']
    out = list(range(int(frag[0]['startRes'])))
    for F in frag:
        out += list(range(int(F['startRes']), int(F['startRes']) + int(F['totalNum'])))
    return [g for g in out if g != exclude]


def env_charges(sysd, exclude=-1):
    res, atoms = sysd['res'], sysd['atoms']
    idx = np.concatenate([np.arange(res[g]['atomStart'], res[g]['atomStart'] + res[g]['atomNum'])
                          for g in present_groups(sysd, exclude)])
    return atoms['position'][idx].astype(np.float64), atoms['charge'][idx].astype(np.float64)


def pose_energy_ewald(sysd, species, centre, pose, beta, env, exclude=-1):
    """FP64 trial energy in Ewald mode: LJ + K q q' erfc(beta r)/r over the
    groups passing calcEnergy's FP32 group-centre test, plus sum_f q_f
    phi_env(x_f) with env an EwaldEnv of every present charge but the
    excluded molecule's."""
    from oracle import oracle as O
    info = sysd['info'][0]
    L = np.asarray(info['cryst'], np.float64)
    Lf = np.asarray(info['cryst'], np.float32)
    rc = np.float32(info['cutoff'])
    Y = int(info['ffYNum'])
    ff = sysd['ff'].astype(np.float64)
    F = sysd['frag'][species]
    nf = int(F['num_atoms'])
    qf = F['atoms']['charge'][:nf].astype(np.float64)
    tf = F['atoms']['type'][:nf]
    P = np.asarray(pose, np.float32).astype(np.float64)
    c32 = np.asarray(centre, np.float32)
    e = 0.0
    res, atoms = sysd['res'], sysd['atoms']
    for g in present_groups(sysd, exclude):
        if not O.distanceP32(c32, res[g]['position'], Lf) <= rc:
            continue
        a0, na = int(res[g]['atomStart']), int(res[g]['atomNum'])
        A = atoms[a0:a0 + na]
        xa = A['position'].astype(np.float64)
        for f in range(nf):
            d = xa - P[f]
            d -= np.where(d / L >= 0, np.floor(d / L + 0.5), np.ceil(d / L - 0.5)) * L
            r2 = (d * d).sum(1)
            r = np.sqrt(r2)
            ti = tf[f] * Y + A['type']
            sig, eps = ff[2 * ti], ff[2 * ti + 1]
            s2 = sig * sig * 100.0 / r2
            s6 = s2 ** 3
            e += float((4 * eps * (s6 * s6 - s6)).sum())
            e += float((KCOUL * qf[f] * A['charge'].astype(np.float64) * erfc(beta * r) / r).sum())
    e += float((qf * env.phi(P)).sum())
    return e
