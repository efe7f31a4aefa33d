"""Synthetic workloads of the benchmark and the large-size tests, built through
the reference-shaped Tree API (so the same construction runs on the HIP
library, the reference harness and the oracle).

octree_poisson: the isolated 3D FAS multigrid solve on a frozen octree —
boxes of 16^3 cells on a 2x2x2 base, uniform level 4 (4096 boxes) plus a
box-aligned sub-volume refined to level 5, 64 seeded Gaussian charge blobs
(widths 0.1-1 mm, charge density +-1e18 e/m^3), phi = V on the bottom face,
phi = 0 on the top face, Neumann sides.
"""
import numpy as np

MASK = (1 << 64) - 1
EPS0 = 8.8541878176e-12   # fosito/m_units_constants.f90:34
ELEC = 1.6022e-19         # fosito/m_units_constants.f90:35


def splitmix64(seed):
    """SplitMix64 stream (Steele, Lea, Flood 2014), as uniform doubles in [0,1)."""
    s = seed & MASK
    while True:
        s = (s + 0x9E3779B97F4A7C15) & MASK
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        z = z ^ (z >> 31)
        yield (z >> 11) * (1.0 / (1 << 53))


# Refinement regions of the developed batched runs, (lvl, lo, hi) in (r, z):
# everything to level 4, a 2 mm wide column to level 6 and a 0.6 mm wide
# channel below the seed to level 8, the mesh of a streamer some 14 mm long
# (about 1,100 boxes of 8 x 8 cells per run)
BATCH_REGIONS = [(4, [0.0, 0.0], [16e-3, 32e-3]),
                 (6, [0.0, 8e-3], [2e-3, 28e-3]),
                 (8, [0.0, 12e-3], [0.6e-3, 26e-3])]


def batch_configs(n, lvls_max, regions=None, init_on_device=False):
    """The batched cylindrical workload: n runs of 16 x 32 mm (1 x 2 base
    boxes of 8 x 8 cells), run k drawn from SplitMix64 seeded 20261203 + k:
    applied field in [1.5, 2.5] MV/m, seed width in [0.2, 0.8] mm, background
    in [1e14, 1e16] m^-3 (log-uniform), the seed on the axis at z = 24 mm.
    regions: refine_regions of every run (None: the reference's criteria
    only)."""
    from . import streamer as ST
    cfgs = []
    for k in range(n):
        g = splitmix64(20261203 + k)
        fld, w, bg = 1.5e6 + 1.0e6 * next(g), 0.2e-3 + 0.6e-3 * next(g), 10 ** (14 + 2 * next(g))
        cfgs.append(ST.Config(variant="cyl", domain_len=16e-3, n_base=(1, 2),
                              lvls_max=lvls_max, applied_fld=fld, seed_width=[w], bg_dens=bg,
                              seed_r0=[[0.0, 24e-3]], seed_r1=[[0.0, 24e-3]],
                              refine_regions=list(regions or []),
                              init_on_device=init_on_device))
    return cfgs


def base_grid(dim, n_base):
    """ix_list / nb_list of an n_base[0] x ... block of base boxes, physical
    boundaries (-1) on the domain faces (Afivo a$D_set_base conventions)."""
    idx = {}
    ix_list = []
    n = 0
    rng = [range(n_base[d]) for d in range(dim)]
    grid = np.stack(np.meshgrid(*rng, indexing="ij"), -1).reshape(-1, dim)
    # i fastest
    order = np.lexsort(grid.T)
    for g in grid[order]:
        n += 1
        idx[tuple(g)] = n
        ix_list.append([int(v) + 1 for v in g])
    nb_list = []
    for g in grid[order]:
        nbs = []
        for d in range(dim):
            for s in (-1, 1):
                q = list(g)
                q[d] += s
                nbs.append(idx.get(tuple(q), -1))
        nb_list.append(nbs)
    return np.array(ix_list, np.int32), np.array(nb_list, np.int32)


class OctreePoisson:
    """Parameters of the 3D octree Poisson workload."""

    def __init__(self, n_cell=16, uniform_lvl=4, fine_lvl=5, domain=8e-3,
                 n_blobs=64, seed=20261203, voltage=16e3, refine_frac=(10, 10, 16),
                 n_cycle_base=8):
        self.dim = 3
        self.n_cell = n_cell
        self.uniform_lvl = uniform_lvl
        self.fine_lvl = fine_lvl
        self.L = domain
        self.n_base = (2, 2, 2)
        self.dr = domain / (self.n_base[0] * n_cell)
        self.n_blobs = n_blobs
        self.seed = seed
        self.voltage = voltage
        # refined sub-volume, in boxes of the uniform level, centred in x, y
        self.refine_boxes = refine_frac
        self.n_cycle_base = n_cycle_base
        # i_phi, i_tmp, i_rhs
        self.ivs = (1, 2, 3)

    def blobs(self):
        u = splitmix64(self.seed)
        amp0 = 1e18 * ELEC / EPS0
        out = []
        for _ in range(self.n_blobs):
            c = [next(u) * self.L for _ in range(3)]
            w = 1e-4 + next(u) * (1e-3 - 1e-4)
            sgn = 1.0 if next(u) < 0.5 else -1.0
            out.append(c + [w, sgn * amp0])
        return np.array(out, float)

    def flags(self, b):
        """Refinement flags by id for the current tree (uniform to
        uniform_lvl, then the sub-volume to fine_lvl)."""
        n = len(b["lvl"])
        f = np.zeros(n, np.int32)
        nbox = self.n_base[0] * 2 ** (self.uniform_lvl - 1)  # boxes per dim
        lo = [(nbox - self.refine_boxes[d]) // 2 for d in range(3)]
        for i in range(n):
            if not b["in_use"][i]:
                continue
            lvl = b["lvl"][i]
            if lvl < self.uniform_lvl:
                f[i] = 1
            elif lvl == self.uniform_lvl and self.fine_lvl > self.uniform_lvl:
                ix = b["ix"][i] - 1
                inside = all(lo[d] <= ix[d] < lo[d] + self.refine_boxes[d] for d in range(3))
                f[i] = 1 if inside else 0
        return f

    def build(self, T):
        ix, nb = base_grid(3, self.n_base)
        T.set_base(ix, nb)
        for _ in range(self.fine_lvl + 1):
            a, r = T.adjust(self.flags(T.boxes()))
            if a == 0 and r == 0:
                break
        return T

    def bc(self):
        # faces lx hx ly hy lz hz: Neumann sides, V at the bottom, 0 at the top
        kinds = np.array([0, 0, 0, 0, 2, 1], np.int32)
        vals = np.array([0, 0, 0, 0, self.voltage, 0], float)
        return kinds, vals

    def leaf_cells(self, T):
        info = T.info()
        n = 0
        for lvl in range(1, info["max_lvl"] + 1):
            n += len(T.level(lvl)[1]) * self.n_cell ** 3
        return n

    def level_boxes(self, T):
        info = T.info()
        return {lvl: len(T.level(lvl)[0]) for lvl in range(info["min_lvl"], info["max_lvl"] + 1)}
