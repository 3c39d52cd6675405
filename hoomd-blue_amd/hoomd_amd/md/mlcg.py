"""Learned coarse-grained (CG) site energy and the tools that build a CG model from a
fine-grained run.

The reference has no learned potential. The force here is the reference's two-pass many-body
form (hoomd/metal/EAMForceCompute.cc:359-560: a density per particle summed over neighbours,
an embedding energy of it, then a pair pass carrying both partners' embedding derivatives) with
16 Gaussian radial-basis densities and an MLP 16 -> 64 -> 64 -> 1 (SiLU) as the embedding:

    G_k(i) = sum_j exp(-(r_ij - mu_k)^2 / (2 w^2)) f_c(r_ij),  f_c(r) = (cos(pi r/r_c) + 1)/2
    E_i    = MLP((G(i) - gmean) * gscale)

It runs on the GPU in three kernels per step (densities, site MLP on bf16 MFMA with FP32
accumulation, forces) and trains by force matching with the project's own backward and Adam
kernels (``SiteEnergyMLP.train``). The cosine cutoff makes every density and its derivative
vanish at r_c.

Workflow (multiscale coarse graining by force matching):

1. ``SiteMap`` groups fine-grained particles into sites (``SiteMap.linear_chains``: consecutive
   beads of each chain, 5:1 for the Kremer-Grest melt).
2. ``FrameSet.record`` maps frames of the fine-grained run on the GPU: site centres of mass and
   the summed conservative forces.
3. ``boltzmann_harmonic`` fits the CG bonds from the mapped bond-length distribution, and
   ``rdf_prior_table`` a short-range repulsive pair prior from the mapped site RDF (a Table
   pair force): the excluded volume of the blobs, which the frames sample too rarely at short
   distances for the network to learn.
4. ``SiteEnergyMLP.train`` matches the remaining force (mapped force minus CG bond and prior
   forces, ``FrameSet.subtract_harmonic`` / ``FrameSet.subtract_pair``).
"""

import ctypes as C

import numpy as np

from .. import _lib
from .force import Force

NK = _lib.MLP_NK
H = 64
NPARAM = _lib.MLP_NPARAM
OFF = {"W1": 0, "b1": H * NK, "W2": H * NK + H, "b2": H * NK + H + H * H,
       "w3": H * NK + 2 * H + H * H, "b3": H * NK + 3 * H + H * H}


def init_params(seed=0):
    """Glorot-uniform weights and zero biases for the 16 -> 64 -> 64 -> 1 network (float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    for name, fan_in, fan_out, n in (("W1", NK, H, H * NK), ("W2", H, H, H * H),
                                     ("w3", H, 1, H)):
        a = np.sqrt(6.0 / (fan_in + fan_out))
        p[OFF[name]:OFF[name] + n] = rng.uniform(-a, a, n)
    return p


class SiteEnergyMLP(Force):
    """SiteEnergyMLP(nlist, r_cut=3.0, width=0.2, mu0=0.0, dmu=0.2, params=None, seed=0)

    The learned site energy as a force of the simulation. ``params`` is the flat float32
    vector W1[64][16], b1[64], W2[64][64], b2[64], w3[64], b3; ``gmean`` and ``gscale``
    normalise the densities. Bonded pairs are left out when the neighbour list excludes
    bonds."""

    _log_namespace = ("md", "mlcg", "SiteEnergyMLP")

    def __init__(self, nlist, r_cut=3.0, width=0.2, mu0=0.0, dmu=0.2, params=None, seed=0):
        super().__init__()
        self.nlist = nlist
        self.r_cut, self.width, self.mu0, self.dmu = float(r_cut), float(width), float(mu0), \
            float(dmu)
        self._params = np.ascontiguousarray(
            init_params(seed) if params is None else params, np.float32).copy()
        if self._params.size != NPARAM:
            raise ValueError(f"params must hold {NPARAM} values")
        self._gmean = np.zeros(NK)
        self._gscale = np.ones(NK)
        self._table = None

    def _model(self):
        m = _lib.HmMlpModel()
        m.r_cut, m.width, m.mu0, m.dmu = self.r_cut, self.width, self.mu0, self.dmu
        for k in range(NK):
            m.gmean[k] = float(self._gmean[k])
            m.gscale[k] = float(self._gscale[k])
        return m

    def _signature(self):
        return ("SiteEnergyMLP", self.r_cut, id(self), self.nlist._signature())

    def _attached(self):
        sim = self._simulation
        return sim is not None and sim._system is not None and self._force_id is not None

    def _attach(self, sim, h):
        fid = _lib.lib().hm_sys_add_mlp_cg(h, C.byref(self._model()), self._params.ctypes.data)
        _lib.check(fid, "mlcg.SiteEnergyMLP")
        self._force_id = fid
        self._simulation = sim
        self.nlist._simulation = sim
        if self._table is not None:
            self._push_table()

    def _pull(self):
        """Keep the trained parameters when the simulation rebuilds its system."""
        if self._attached():
            self.params  # the getter downloads them into the host copy

    def _push_table(self):
        t = self._table
        if t is None:
            r = _lib.lib().hm_sys_mlp_set_pair_table(self._simulation._system, self._force_id, 0,
                                                     0.0, None, None)
        else:
            r = _lib.lib().hm_sys_mlp_set_pair_table(
                self._simulation._system, self._force_id, len(t["U"]), float(t["r_min"]),
                t["U"].ctypes.data, t["F"].ctypes.data)
        _lib.check(r, "mlcg.SiteEnergyMLP.pair_table")

    @property
    def pair_table(self):
        """Optional tabulated pair term (dict r_min, U, F on [r_min, r_cut), the form of
        md.pair.Table params) evaluated in the site energy's own force pass over the same pairs:
        a structural prior at no extra pass. None: no pair term."""
        if self._table is None:
            return None
        return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self._table.items()}

    @pair_table.setter
    def pair_table(self, value):
        if value is None:
            self._table = None
        else:
            U = np.ascontiguousarray(value["U"], np.float64).ravel().copy()
            F = np.ascontiguousarray(value["F"], np.float64).ravel().copy()
            if U.size == 0 or U.size != F.size:
                raise ValueError("pair_table: U and F need the same, nonzero length")
            self._table = dict(r_min=float(value.get("r_min", 0.0)), U=U, F=F)
        if self._attached():
            self._push_table()

    # ------------------------------------------------------------------ parameters
    @property
    def params(self):
        if self._attached():
            _lib.check(_lib.lib().hm_sys_mlp_get_params(self._simulation._system, self._force_id,
                                                        self._params.ctypes.data), "params")
        return self._params.copy()

    @params.setter
    def params(self, value):
        v = np.ascontiguousarray(value, np.float32).ravel()
        if v.size != NPARAM:
            raise ValueError(f"params must hold {NPARAM} values")
        self._params = v.copy()
        if self._attached():
            _lib.check(_lib.lib().hm_sys_mlp_set_params(self._simulation._system,
                                                        self._force_id, self._params.ctypes.data),
                       "params")

    @property
    def normalization(self):
        return self._gmean.copy(), self._gscale.copy()

    @normalization.setter
    def normalization(self, value):
        gmean, gscale = value
        self._gmean = np.asarray(gmean, np.float64).reshape(NK).copy()
        self._gscale = np.asarray(gscale, np.float64).reshape(NK).copy()
        if self._attached():
            _lib.check(_lib.lib().hm_sys_mlp_set_model(self._simulation._system, self._force_id,
                                                       C.byref(self._model())), "normalization")

    def descriptors(self, frames=None, k=0):
        """The densities G, [N, 16] by tag, of the current state or of frame k of a FrameSet
        (whose positions then replace the state's)."""
        sim = self._simulation
        sim._attach()
        G = np.zeros((sim.state.N_particles, NK), np.float32)
        fp = frames.ptr_pos(k) if frames is not None else None
        _lib.check(_lib.lib().hm_sys_mlp_descriptors(sim._system, self._force_id, fp,
                                                     G.ctypes.data), "descriptors")
        return G

    def fit_normalization(self, G):
        """gmean, gscale from densities G [..., 16]: zero mean and unit spread per feature."""
        G = np.asarray(G, np.float64).reshape(-1, NK)
        self.normalization = (G.mean(axis=0), 1.0 / (G.std(axis=0) + 1e-3))

    # ------------------------------------------------------------------ training
    def train(self, frames, steps, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, order=None,
              target=None, frames_per_step=1):
        """Force matching on device frames: Adam steps over frames (in ``order``, default
        cycling), each on loss = sum_i |F_i - T_i|^2 / (3N), the gradient averaged over
        ``frames_per_step`` frames (``order`` then lists steps * frames_per_step frames).
        ``target`` is a FrameSet of target forces when they differ from ``frames`` (e.g. mapped
        minus bond forces). Returns the loss of every step, taken before its update."""
        sim = self._simulation
        sim._attach()
        t = target if target is not None else frames
        a = _lib.HmMlpTrainArgs()
        a.nframes = frames.nframes
        a.frame_pos = frames.d_pos.ptr
        a.frame_force = t.d_force.ptr
        a.nsteps = int(steps)
        a.frames_per_step = max(1, int(frames_per_step))
        ordr = None
        if order is not None:
            ordr = np.ascontiguousarray(order, np.uint32)
            if ordr.size < steps * a.frames_per_step:
                raise ValueError("order needs frames_per_step frames per step")
            a.order = ordr.ctypes.data
        a.lr, a.beta1, a.beta2, a.eps = float(lr), float(betas[0]), float(betas[1]), float(eps)
        loss = np.zeros(max(int(steps), 1))
        a.loss = loss.ctypes.data
        _lib.check(_lib.lib().hm_sys_mlp_train(sim._system, self._force_id, C.byref(a)), "train")
        return loss[:int(steps)]

    def loss_grad(self, frames, k=0, target=None):
        """(loss, gradient [NPARAM]) of frame k, no update."""
        sim = self._simulation
        sim._attach()
        t = target if target is not None else frames
        loss = C.c_double()
        g = np.zeros(NPARAM, np.float32)
        _lib.check(_lib.lib().hm_sys_mlp_loss_grad(
            sim._system, self._force_id, frames.ptr_pos(k), t.ptr_force(k), C.byref(loss),
            g.ctypes.data), "loss_grad")
        return loss.value, g


class SiteMap:
    """Sites as groups of particle tags (CSR: site I holds tags[row[I]:row[I+1]])."""

    def __init__(self, row, tags):
        self.row = np.ascontiguousarray(row, np.uint32)
        self.tags = np.ascontiguousarray(tags, np.uint32)
        self.nsites = len(self.row) - 1
        self._chain = None

    @classmethod
    def linear_chains(cls, n_chains, chain_len, per_site):
        """Consecutive groups of per_site beads along each chain (tags chain * chain_len + b)."""
        if chain_len % per_site:
            raise ValueError("chain_len must be a multiple of per_site")
        nsites = n_chains * (chain_len // per_site)
        m = cls(np.arange(nsites + 1) * per_site, np.arange(n_chains * chain_len))
        m._chain = (n_chains, chain_len // per_site)
        return m

    def bonds(self):
        """CG bonds between consecutive sites of each chain, [M, 2] site indices."""
        if self._chain is None:
            raise ValueError("bonds() needs a SiteMap.linear_chains map")
        nc, ns = self._chain
        first = (np.arange(nc)[:, None] * ns + np.arange(ns - 1)[None, :]).ravel()
        return np.stack([first, first + 1], axis=1).astype(np.uint32)

    def attach(self, sim):
        sim._attach()
        _lib.check(_lib.lib().hm_sys_set_site_map(sim._system, self.nsites, self.row.ctypes.data,
                                                  self.tags.ctypes.data), "site map")


class FrameSet:
    """Device-resident frames of nsites sites: FP32 positions and forces, [nframes][nsites][3]
    each, in site (tag) order."""

    def __init__(self, nframes, nsites):
        self.nframes, self.nsites = int(nframes), int(nsites)
        self.d_pos = _lib.DeviceArray((self.nframes * self.nsites * 3,), np.float32, fill=0)
        self.d_force = _lib.DeviceArray((self.nframes * self.nsites * 3,), np.float32, fill=0)

    def _off(self, k):
        k = int(k)
        if not 0 <= k < self.nframes:
            raise IndexError(k)
        return 4 * 3 * self.nsites * k

    def ptr_pos(self, k):
        return self.d_pos.ptr + self._off(k)

    def ptr_force(self, k):
        return self.d_force.ptr + self._off(k)

    def record(self, sim, k):
        """Map the simulation's current state into frame k (the site map must be attached)."""
        _lib.check(_lib.lib().hm_sys_map_frame(sim._system, self.ptr_pos(k), self.ptr_force(k)),
                   "map frame")

    def _get(self, ptr):
        out = np.zeros((self.nsites, 3), np.float32)
        _lib.check(_lib.lib().hm_memcpy_dtoh(out.ctypes.data, ptr, out.nbytes), "download")
        return out

    def positions(self, k):
        return self._get(self.ptr_pos(k))

    def forces(self, k):
        return self._get(self.ptr_force(k))

    def subtract_harmonic(self, bonds, k, r0, L):
        """Replace every frame's forces by the forces minus those of harmonic bonds (k, r0)
        between the site pairs ``bonds`` [M, 2] (on the device; cubic box edge L)."""
        b = np.asarray(bonds, np.int64).reshape(-1, 2)
        ends = np.concatenate([b, b[:, ::-1]])
        ends = ends[np.lexsort((ends[:, 1], ends[:, 0]))]
        row = np.searchsorted(ends[:, 0], np.arange(self.nsites + 1)).astype(np.uint32)
        partner = np.ascontiguousarray(ends[:, 1], np.uint32)
        d_row = _lib.DeviceArray.from_numpy(row)
        d_partner = _lib.DeviceArray.from_numpy(partner if len(partner) else np.zeros(1, np.uint32))
        _lib.check(_lib.lib().hm_frames_sub_harmonic(
            self.nframes, self.nsites, self.d_pos.ptr, self.d_force.ptr, self.d_force.ptr,
            d_row.ptr, d_partner.ptr, float(k), float(r0), _lib.box(L), None), "bond forces")
        _lib.check(_lib.lib().hm_synchronize(), "bond forces")

    def subtract_pair(self, sim, force):
        """Replace every frame's forces by the forces minus those of the pair force ``force``
        (LJ or Table, attached to ``sim``, whose sites are these frames' sites) at the frame's
        positions. The simulation's positions are left at the last frame."""
        sim._attach()
        for k in range(self.nframes):
            _lib.check(_lib.lib().hm_sys_frame_subtract_force(
                sim._system, force._force_id, self.ptr_pos(k), self.ptr_force(k)), "prior force")

    def set(self, k, positions=None, forces=None):
        for arr, ptr in ((positions, self.ptr_pos(k)), (forces, self.ptr_force(k))):
            if arr is not None:
                a = np.ascontiguousarray(arr, np.float32).reshape(self.nsites, 3)
                _lib.check(_lib.lib().hm_memcpy_htod(ptr, a.ctypes.data, a.nbytes), "upload")


# ---------------------------------------------------------------------- analysis
def _wrap(d, L):
    return d - L * np.rint(d / L)


def bond_lengths(pos, L, bonds):
    d = _wrap(pos[bonds[:, 0]] - pos[bonds[:, 1]], L)
    return np.sqrt((d * d).sum(axis=1))


def harmonic_bond_forces(pos, L, bonds, k, r0):
    """Forces of U = k (r - r0)^2 / 2 on every site, [N, 3]."""
    d = _wrap(pos[bonds[:, 0]] - pos[bonds[:, 1]], L)
    r = np.sqrt((d * d).sum(axis=1))
    fdr = k * (r0 / r - 1.0)
    f = d * fdr[:, None]
    out = np.zeros_like(pos, dtype=np.float64)
    np.add.at(out, bonds[:, 0], f)
    np.add.at(out, bonds[:, 1], -f)
    return out


def boltzmann_harmonic(r, kT=1.0, nbins=80, min_count=20):
    """Harmonic bond (k, r0) fitted to the Boltzmann inversion U(r) = -kT ln(P(r) / r^2) of
    bond-length samples r: a least-squares parabola over the bins holding at least min_count
    samples, weighted by their counts."""
    r = np.asarray(r, np.float64)
    h, e = np.histogram(r, bins=nbins)
    c = 0.5 * (e[1:] + e[:-1])
    ok = h >= min_count
    U = -kT * np.log(h[ok] / (c[ok] ** 2))
    a, b, _ = np.polyfit(c[ok], U, 2, w=np.sqrt(h[ok]))
    if a <= 0:
        raise RuntimeError("Boltzmann inversion: the bond distribution is not single-peaked")
    return 2.0 * a, -b / (2.0 * a)


def rdf(pos, L, r_max, nbins=60, exclude_bonded=None):
    """Radial distribution function of one frame: (bin centres, g(r)). exclude_bonded (pairs
    [M, 2]) are removed from the counts."""
    from scipy.spatial import cKDTree
    pos = np.asarray(pos, np.float64)
    N = len(pos)
    x = np.mod(pos + L / 2, L)
    x[x >= L] = 0.0
    tree = cKDTree(x, boxsize=L)
    edges = np.linspace(0.0, r_max, nbins + 1)
    pairs = tree.query_pairs(r_max, output_type="ndarray")
    d = _wrap(x[pairs[:, 0]] - x[pairs[:, 1]], L)
    h, _ = np.histogram(np.sqrt((d * d).sum(axis=1)), bins=edges)
    shell = 2.0 * h.astype(np.float64)  # ordered pairs
    if exclude_bonded is not None and len(exclude_bonded):
        rb = bond_lengths(pos, L, np.asarray(exclude_bonded))
        hb, _ = np.histogram(rb, bins=edges)
        shell -= 2.0 * hb
    rho = N / L ** 3
    vol = 4.0 / 3.0 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    return 0.5 * (edges[1:] + edges[:-1]), shell / (N * rho * vol)


def rdf_prior_table(r, g, kT=1.0, g_join=0.9, width=400, min_count_g=0.02):
    """Repulsive prior from the short-range Boltzmann inversion of a site RDF (the first step of
    iterative Boltzmann inversion): U(r) = -kT ln g(r) up to the first bin where g reaches
    g_join (its r is the returned cut-off), shifted to zero there. Below the first bin with
    g >= min_count_g the samples are too few, so U continues along a straight line with the
    slope of the first well-sampled bins (a finite repulsive wall). Returns (Table params,
    r_cut)."""
    r = np.asarray(r, np.float64)
    g = np.asarray(g, np.float64)
    above = np.nonzero(g >= g_join)[0]
    if len(above) == 0:
        raise ValueError("rdf_prior_table: g(r) never reaches g_join")
    ic = int(above[0])
    rc = float(r[ic])
    ok = np.nonzero(g[:ic + 1] >= min_count_g)[0]
    i0 = int(ok[0])
    U = np.full(len(r), np.nan)
    U[i0:ic + 1] = -kT * np.log(g[i0:ic + 1])
    U -= U[ic]
    i1 = min(i0 + 3, ic)
    slope = (U[i1] - U[i0]) / (r[i1] - r[i0]) if i1 > i0 else -10.0 * kT
    slope = min(slope, -kT)  # a wall: U falls with r
    U[:i0] = U[i0] + slope * (r[:i0] - r[i0])
    # the wall continues down to r = 0
    r_ext = np.concatenate([[0.0], r[:ic + 1]])
    U_ext = np.concatenate([[U[i0] + slope * (0.0 - r[i0])], U[:ic + 1]])
    rr = np.linspace(0.0, rc, width, endpoint=False)
    Ut = np.interp(rr, r_ext, U_ext)
    Ft = -np.gradient(Ut, rr)
    return dict(r_min=0.0, U=Ut, F=Ft), rc


def regrid_table(params, r_cut, r_cut_new, width=None):
    """Table params (r_min 0 on [0, r_cut)) carried onto [0, r_cut_new) with U = 0 beyond the
    old cut-off: the same forces, on the grid an iterative refinement can reshape out to
    r_cut_new."""
    U = np.asarray(params["U"], np.float64)
    w = len(U) if width is None else int(width)
    r_old = np.linspace(0.0, r_cut, len(U), endpoint=False)
    r = np.linspace(0.0, r_cut_new, w, endpoint=False)
    Un = np.interp(r, np.concatenate([r_old, [r_cut]]), np.concatenate([U, [0.0]]), right=0.0)
    return dict(r_min=0.0, U=Un, F=-np.gradient(Un, r))


def ibi_update(params, r_cut, r, g_cur, g_target, kT=1.0, damping=0.3, g_min=0.02, taper=0.3):
    """One iterative Boltzmann inversion step on a Table (Reith, Puetz and Mueller-Plathe,
    J. Comput. Chem. 24, 1624 (2003)): U <- U + damping kT ln(g_cur / g_target), the
    correction taken on the RDF bins r where both RDFs are at least g_min, smoothed over three
    bins, held at its first sampled value below them, and brought to zero at r_cut by a cosine
    taper over the last `taper` of the range (so U(r_cut) stays 0). F = -dU/dr."""
    U = np.asarray(params["U"], np.float64)
    rt = np.linspace(0.0, r_cut, len(U), endpoint=False)
    r = np.asarray(r, np.float64)
    ok = (np.asarray(g_cur) >= g_min) & (np.asarray(g_target) >= g_min) & (r < r_cut)
    if ok.sum() < 3:
        raise ValueError("ibi_update: too few sampled bins")
    d = damping * kT * np.log(np.asarray(g_cur)[ok] / np.asarray(g_target)[ok])
    d = np.convolve(np.pad(d, 1, mode="edge"), np.ones(3) / 3.0, mode="valid")
    corr = np.interp(rt, r[ok], d)  # held constant outside the sampled bins
    t = np.clip((r_cut - rt) / taper, 0.0, 1.0)
    corr *= 0.5 * (1.0 - np.cos(np.pi * t))
    Un = U + corr
    return dict(r_min=0.0, U=Un, F=-np.gradient(Un, rt))


def pressure_ramp(params, r_cut, dp, kT=1.0, scale=0.1, dp_scale=1.0):
    """The linear pressure correction used beside iterative Boltzmann inversion (Reith, Puetz
    and Mueller-Plathe 2003): U <- U + A (1 - r / r_cut), A = -scale kT clip(dp / dp_scale, -1,
    1), dp = P_model - P_target. A model below its target pressure gets a repulsive ramp; the
    ramp is zero at r_cut. F = -dU/dr."""
    U = np.asarray(params["U"], np.float64)
    rt = np.linspace(float(params.get("r_min", 0.0)), r_cut, len(U), endpoint=False)
    A = -scale * kT * float(np.clip(dp / dp_scale, -1.0, 1.0))
    Un = U + A * (1.0 - rt / r_cut)
    return dict(r_min=float(params.get("r_min", 0.0)), U=Un, F=-np.gradient(Un, rt))


def soft_cosine_table(A, r_cut, width=200):
    """Table params of the push-off potential U = A (1 + cos(pi r / r_cut)) for r < r_cut, a
    finite repulsion at r = 0 that lets overlapping beads separate."""
    r = np.linspace(0.0, r_cut, width, endpoint=False)
    U = A * (1.0 + np.cos(np.pi * r / r_cut))
    F = A * np.pi / r_cut * np.sin(np.pi * r / r_cut)
    return dict(r_min=0.0, U=U, F=F)
