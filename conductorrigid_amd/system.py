"""HOOMD-shaped system data for the ConductorRigid force compute, and a seeded
generator of rigid conductor clusters.

The reference plugin reads its inputs from HOOMD 1.x (cppmodule/ConductorRigid.cc:420-487):
particle positions/types by global index, particle tags, the user's per-tag
isolated-body map ``body`` and per-body net ``charge`` lists
(cppmodule/ConductorRigid.cc:93-105), and HOOMD's RigidData (centre of mass, body
size, member indices with pitch Nmax, per-particle rigid index).  HOOMD is not
part of this package, so :class:`SystemDefinition` holds exactly those arrays
with the same meaning and layout.

Isolated-body tags follow the reference's convention: every rigid body and every
bead that belongs to no rigid body is one "isolated body" with its own tag and
its own net charge; rigid body ``r`` has tag ``r`` and isolated beads take the
tags after the rigid bodies.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

NOT_MEMBER = 0xFFFFFFFF


@dataclass
class SystemDefinition:
    L: np.ndarray                 # (3,) box lengths; box spans [-L/2, L/2)
    pos: np.ndarray               # (N, 4) float64  x, y, z, type  by global index
    tag: np.ndarray               # (N,) uint32     particle tag of each global index
    body_of_tag: np.ndarray       # (N,) uint32     isolated-body tag per particle tag
    body_charge: np.ndarray       # (N_isolated,) float64 net charge per isolated body
    rigid_membership: np.ndarray  # (N,) uint32     rigid index per global index or NOT_MEMBER
    rigid_members: np.ndarray     # (N_rigid*Nmax,) uint32 global indices, NOT_MEMBER padded
    rigid_size: np.ndarray        # (N_rigid,) uint32
    rigid_com: np.ndarray         # (N_rigid, 4) float64 centre of mass (wrapped)
    Nmax: int
    meta: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return int(self.pos.shape[0])

    @property
    def N_rigid(self) -> int:
        return int(self.rigid_size.shape[0])

    @property
    def N_isolated(self) -> int:
        return int(self.body_charge.shape[0])

    def iso_membership(self) -> np.ndarray:
        """Isolated-body tag of each bead by global index (update_membership,
        cppmodule/ConductorRigid.cu:117-141)."""
        return self.body_of_tag[self.tag]

    def body_counts(self) -> np.ndarray:
        return np.bincount(self.iso_membership(), minlength=self.N_isolated)

    def validate(self) -> None:
        N = self.N
        if self.pos.shape != (N, 4) or self.pos.dtype != np.float64:
            raise ValueError("pos must be float64 (N,4)")
        if sorted(self.tag.tolist()) != list(range(N)):
            raise ValueError("tag must be a permutation of 0..N-1")
        iso = self.iso_membership()
        if iso.size and iso.max() >= self.N_isolated:
            raise ValueError("body tag out of range")
        for r in range(self.N_rigid):
            n = int(self.rigid_size[r])
            mem = self.rigid_members[r * self.Nmax:r * self.Nmax + n]
            if np.any(mem == NOT_MEMBER) or np.any(self.rigid_membership[mem] != r):
                raise ValueError(f"rigid body {r} members inconsistent")
        half = 0.5 * self.L
        if np.any(self.pos[:, :3] < -half) or np.any(self.pos[:, :3] >= half):
            raise ValueError("positions must be wrapped into [-L/2, L/2)")


def box_length(n_beads: int, phi: float, radius: float = 1.0) -> float:
    """Cubic box edge for ``n_beads`` spheres of ``radius`` at volume fraction ``phi``."""
    return (n_beads * 4.0 / 3.0 * math.pi * radius ** 3 / phi) ** (1.0 / 3.0)


def _random_rotations(rng: np.random.Generator, m: int) -> np.ndarray:
    q = rng.normal(size=(m, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((m, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _wrap(x: np.ndarray, L: np.ndarray) -> np.ndarray:
    y = x - L * np.floor((x + 0.5 * L) / L)
    # guard the half-open interval against rounding up to +L/2
    y = np.where(y >= 0.5 * L, y - L, y)
    return y


def _body_means(x: np.ndarray, offsets: np.ndarray, bodies: np.ndarray, count: np.ndarray) -> np.ndarray:
    """Mean of the first count[b] rows of each body's slice of x."""
    k = count[bodies]
    kmax = int(k.max())
    cols = np.arange(kmax)
    rows = offsets[bodies, None] + np.minimum(cols[None, :], k[:, None] - 1)
    mask = (cols[None, :] < k[:, None]).astype(np.float64)
    return (x[rows] * mask[:, :, None]).sum(axis=1) / k[:, None]


def _fibonacci_sphere(n: int) -> np.ndarray:
    """n nearly uniform unit vectors (golden-angle spiral)."""
    k = np.arange(n) + 0.5
    z = 1.0 - 2.0 * k / n
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    t = math.pi * (3.0 - math.sqrt(5.0)) * k
    return np.stack([r * np.cos(t), r * np.sin(t), z], axis=1)


def _grow_exhaustive(unwrapped: np.ndarray, placed: np.ndarray, offsets: np.ndarray, count: np.ndarray,
                     sizes: np.ndarray, L: np.ndarray, tol: float, n_dir: int = 384) -> int:
    """Grow each unfinished body by up to its missing beads, one bead at a time, at the free
    touching site (distance 2 to a bead of its own body, >= 2 from every bead) closest to
    its centre, trying n_dir directions around every one of its beads.  Returns the number
    of beads placed."""
    from scipy.spatial import cKDTree

    Lval = float(L[0])
    dirs = 2.0 * _fibonacci_sphere(n_dir)
    placed_total = 0
    for b in np.nonzero(count < sizes)[0]:
        while count[b] < sizes[b]:
            idx_placed = np.nonzero(placed)[0]
            tree = cKDTree(_wrap(unwrapped[idx_placed], L) + 0.5 * L, boxsize=Lval)
            own = unwrapped[offsets[b]:offsets[b] + count[b]]
            cand = (own[:, None, :] + dirs[None, :, :]).reshape(-1, 3)
            dist, _ = tree.query(_wrap(cand, L) + 0.5 * L, k=1)
            ok = dist >= 2.0 - tol
            if not np.any(ok):
                break
            cand = cand[ok]
            c = cand[np.argmin(np.linalg.norm(cand - own.mean(axis=0), axis=1))]
            slot = offsets[b] + count[b]
            unwrapped[slot] = c
            placed[slot] = True
            count[b] += 1
            placed_total += 1
    return placed_total


def make_system(n_bodies: int, body_sizes, phi: float, seed: int, charges: str = "alternating",
                n_isolated_beads: int = 0, spatial_sort: bool = True,
                n_candidates: int = 16, dense_growth: bool = False) -> SystemDefinition:
    """Seeded random system of rigid conductor clusters (plus optional isolated beads).

    ``body_sizes`` is an int (all bodies that size) or a sequence to draw from.
    Every body starts as one seed bead on a random site of an FCC lattice (with
    jitter) and all bodies then grow together, one bead per body per round:
    a new bead touches (centre distance 2) a random bead of its own body,
    overlaps no bead of any body (periodic minimum image), and of
    ``n_candidates`` proposals the one closest to the body's centre is kept,
    which makes compact random aggregates.  When every unfinished body fails its
    proposals for several rounds, the caged bodies are re-seeded at free spots; with
    ``dense_growth`` a search over every bead of each unfinished body and 384 directions
    for a free touching site comes first (this reaches volume fraction 0.35 with 8-, 20-
    and 40-bead bodies, where re-seeding stalls above ~0.33).  ``charges`` is "alternating"
    (+1/-1 by body), "uniform" (U[-2, 2] shifted to total 0) or "zero".
    """
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    if np.isscalar(body_sizes):
        sizes = np.full(n_bodies, int(body_sizes), dtype=np.int64)
    else:
        sizes = rng.choice(np.asarray(body_sizes, dtype=np.int64), size=n_bodies)
    n_beads = int(sizes.sum()) + n_isolated_beads
    Lval = box_length(n_beads, phi)
    L = np.array([Lval] * 3)

    n_sites_needed = n_bodies + n_isolated_beads
    m = max(1, int(math.ceil((n_sites_needed / 4.0) ** (1.0 / 3.0))))
    a = Lval / m
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    sites = ((cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) + 0.25) * a - 0.5 * Lval
    centres = sites[rng.permutation(sites.shape[0])[:n_sites_needed]]
    centres = centres + rng.uniform(-0.05, 0.05, size=centres.shape) * a

    offsets = np.concatenate([[0], np.cumsum(sizes)])
    owner = np.concatenate([np.repeat(np.arange(n_bodies), sizes), n_bodies + np.arange(n_isolated_beads)])
    unwrapped = np.zeros((n_beads, 3))
    placed = np.zeros(n_beads, dtype=bool)
    unwrapped[offsets[:-1]] = centres[:n_bodies]
    placed[offsets[:-1]] = True
    if n_isolated_beads:
        unwrapped[offsets[-1]:] = centres[n_bodies:]
        placed[offsets[-1]:] = True
    count = np.ones(n_bodies, dtype=np.int64)
    tol = 1e-9
    stall = 0
    restarts = 0
    # placed beads live in a large tree rebuilt only after many insertions, plus a
    # small tree of the beads placed since (rebuilt every round)
    big_tree = None
    recent: list[np.ndarray] = []
    n_recent = 0
    while np.any(count < sizes):
        active = np.nonzero(count < sizes)[0]
        if big_tree is None or n_recent > 20000:
            idx_placed = np.nonzero(placed)[0]
            big_tree = cKDTree(_wrap(unwrapped[idx_placed], L) + 0.5 * L, boxsize=Lval)
            recent, n_recent = [], 0
        small_tree = cKDTree(_wrap(np.concatenate(recent), L) + 0.5 * L, boxsize=Lval) if recent else None
        A, K = active.size, n_candidates
        base_k = (rng.random((A, K)) * count[active, None]).astype(np.int64)
        base = unwrapped[offsets[active, None] + base_k]
        d = rng.normal(size=(A, K, 3))
        d /= np.linalg.norm(d, axis=2, keepdims=True)
        cand = base + 2.0 * d
        cq = _wrap(cand.reshape(-1, 3), L) + 0.5 * L
        dist, _ = big_tree.query(cq, k=1)
        if small_tree is not None:
            dist = np.minimum(dist, small_tree.query(cq, k=1)[0])
        valid = (dist.reshape(A, K) >= 2.0 - tol)
        com = np.stack([unwrapped[offsets[b]:offsets[b] + count[b]].mean(axis=0) for b in active]) \
            if A < 2000 else _body_means(unwrapped, offsets, active, count)
        score = np.linalg.norm(cand - com[:, None, :], axis=2)
        score[~valid] = np.inf
        best = np.argmin(score, axis=1)
        ok = np.isfinite(score[np.arange(A), best])
        chosen_b = active[ok]
        chosen = cand[np.arange(A)[ok], best[ok]]
        if chosen_b.size:
            t2 = cKDTree(_wrap(chosen, L) + 0.5 * L, boxsize=Lval)
            pr = t2.query_pairs(2.0 - tol, output_type="ndarray")
            drop = np.zeros(chosen_b.size, dtype=bool)
            if pr.size:
                drop[np.maximum(pr[:, 0], pr[:, 1])] = True
            chosen_b, chosen = chosen_b[~drop], chosen[~drop]
            slot = offsets[chosen_b] + count[chosen_b]
            unwrapped[slot] = chosen
            placed[slot] = True
            count[chosen_b] += 1
            recent.append(chosen.copy())
            n_recent += chosen.shape[0]
        stall = 0 if chosen_b.size else stall + 1
        if stall > 8:
            if dense_growth and _grow_exhaustive(unwrapped, placed, offsets, count, sizes, L, tol):
                # random proposals found no room, a search over every bead of each unfinished
                # body and a dense set of directions did
                big_tree, recent, n_recent = None, [], 0
                stall = 0
                continue
            # every remaining body is caged in: re-seed each at a random free spot
            restarts += 1
            if restarts > 200:
                raise RuntimeError("cluster growth stalled; lower the volume fraction")
            caged = np.nonzero(count < sizes)[0]
            for b in caged:
                placed[offsets[b]:offsets[b + 1]] = False
            idx_placed = np.nonzero(placed)[0]
            tree = cKDTree(_wrap(unwrapped[idx_placed], L) + 0.5 * L, boxsize=Lval)
            trial = rng.uniform(-0.5, 0.5, size=(max(4096, 64 * caged.size), 3)) * Lval
            room = tree.query(trial + 0.5 * L, k=1)[0]
            order = np.argsort(-room)
            chosen = []
            for t in order:
                if room[t] < 2.0 or len(chosen) == caged.size:
                    break
                p0 = trial[t]
                if all(np.linalg.norm(_wrap(p0 - c, L)) >= 2.0 for c in chosen):
                    chosen.append(p0)
            if len(chosen) < caged.size:
                raise RuntimeError("no free space to re-seed caged bodies; lower the volume fraction")
            for b, p0 in zip(caged, chosen):
                unwrapped[offsets[b]] = p0
                placed[offsets[b]] = True
                count[b] = 1
            big_tree = None
            stall = 0

    com_unwrapped = np.array([unwrapped[offsets[b]:offsets[b + 1]].mean(axis=0) for b in range(n_bodies)]) \
        if n_bodies else np.zeros((0, 3))
    pos_tagorder = _wrap(unwrapped, L)

    # global index order: spatial (cell) sort like HOOMD's particle sorter
    if spatial_sort:
        nb = max(1, int(Lval // 2.0))
        cell = np.floor((pos_tagorder + 0.5 * L) / L * nb).astype(np.int64).clip(0, nb - 1)
        key = (cell[:, 0] * nb + cell[:, 1]) * nb + cell[:, 2]
        order = np.argsort(key, kind="stable")          # order[idx] = tag
    else:
        order = np.arange(n_beads)
    tag = order.astype(np.uint32)
    rtag = np.empty(n_beads, dtype=np.int64)
    rtag[order] = np.arange(n_beads)

    pos = np.zeros((n_beads, 4))
    pos[:, :3] = pos_tagorder[order]

    N_iso = n_bodies + n_isolated_beads
    body_of_tag = owner.astype(np.uint32)          # rigid body r -> tag r; isolated beads after
    Nmax = int(sizes.max()) if n_bodies else 1
    rigid_members = np.full(n_bodies * Nmax, NOT_MEMBER, dtype=np.uint32)
    rigid_membership = np.full(n_beads, NOT_MEMBER, dtype=np.uint32)
    for b in range(n_bodies):
        idxs = rtag[offsets[b]:offsets[b + 1]]
        rigid_members[b * Nmax:b * Nmax + sizes[b]] = idxs
        rigid_membership[idxs] = b
    rigid_com = np.zeros((n_bodies, 4))
    rigid_com[:, :3] = _wrap(com_unwrapped, L) if n_bodies else 0

    if charges == "alternating":
        body_charge = np.where(np.arange(N_iso) % 2 == 0, 1.0, -1.0)
        if N_iso % 2:
            body_charge[-1] = 0.0
    elif charges == "uniform":
        body_charge = rng.uniform(-2.0, 2.0, size=N_iso)
        body_charge -= body_charge.mean()
    elif charges == "zero":
        body_charge = np.zeros(N_iso)
    else:
        raise ValueError(charges)

    sysdef = SystemDefinition(L=L, pos=pos, tag=tag, body_of_tag=body_of_tag, body_charge=body_charge,
                              rigid_membership=rigid_membership, rigid_members=rigid_members,
                              rigid_size=sizes.astype(np.uint32), rigid_com=rigid_com, Nmax=Nmax,
                              meta=dict(seed=seed, phi=phi, n_bodies=n_bodies, charges=charges))
    return sysdef


# Workloads used by the parity tests and the benchmark.
WORKLOADS = {
    # N=1000, 50 bodies x 20 beads, phi=0.2, Q=+-1 alternating, mesh 64^3, P=8, rc=4
    "n1k": dict(n_bodies=50, body_sizes=20, phi=0.2, charges="alternating", seed=1,
                xi=0.929, errortol=1e-6, mesh=(64, 64, 64), P=8, rc=4.0, restart=10, tol=1e-6),
    # N=100k, 5000 x 20, phi=0.3, Q~U[-2,2] neutral, mesh 128^3, P=8, rc=4, restart 30
    "n100k": dict(n_bodies=5000, body_sizes=20, phi=0.3, charges="uniform", seed=2,
                  xi=0.6, errortol=1e-6, mesh=(128, 128, 128), P=8, rc=4.0, restart=30, tol=1e-6),
    # N=1M, 50k x 20, phi=0.3, mesh 256^3, P=8, rc=4, restart 30
    "n1m": dict(n_bodies=50000, body_sizes=20, phi=0.3, charges="uniform", seed=3,
                xi=0.6, errortol=1e-6, mesh=(256, 256, 256), P=8, rc=4.0, restart=30, tol=1e-6),
}


def translate_bodies(s: SystemDefinition, step: float, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """Rigid random translation of every isolated body by a displacement uniform in
    [-step, step]^3 (a stand-in for one dynamics step between force evaluations).
    Returns wrapped bead positions (N, 4) and rigid centres of mass (N_rigid, 4)."""
    rng = np.random.default_rng(seed)
    d = rng.uniform(-step, step, size=(s.N_isolated, 3))
    pos = s.pos.copy()
    pos[:, :3] = _wrap(pos[:, :3] + d[s.iso_membership()], s.L)
    com = s.rigid_com.copy()
    if s.N_rigid:
        # rigid body r carries isolated-body tag r
        com[:, :3] = _wrap(com[:, :3] + d[:s.N_rigid], s.L)
    return pos, com


def make_workload(name: str) -> tuple[SystemDefinition, dict]:
    w = dict(WORKLOADS[name])
    sysdef = make_system(w.pop("n_bodies"), w.pop("body_sizes"), w.pop("phi"), seed=w.pop("seed"),
                         charges=w.pop("charges"))
    return sysdef, w


# Mesh sizes of batched boxes: the lengths the register-resident FFT passes take
# (n = A * B, A = B or 2B), all with binned tilings of tile edge >= P - 1 = 8.
BATCH_MESHES = (32, 64, 128)


def batch_box_params(L: float, errortol: float = 1e-6, h_target: float = 1.0) -> dict:
    """Ewald parameters of one batched box for ``errortol``: the mesh is the smallest
    entry of BATCH_MESHES with spacing <= h_target (capped at 128); xi is the largest
    splitting for which SetParams' mesh rule, N >= 1 + L kcut / pi with
    kcut = 2 xi sqrt(-ln tol) (cppmodule/ConductorRigid.cc:124-129), still fits that
    mesh, and rc = sqrt(-ln tol) / xi (:123).  P = ceil(-2 ln tol / pi) (:201); the
    Gaussian width eta = P xi^2 h^2 / pi is then about 0.5 < 1 for every box."""
    m = BATCH_MESHES[-1]
    for cand in BATCH_MESHES:
        if L / cand <= h_target:
            m = cand
            break
    s = math.sqrt(-math.log(errortol))
    xi = math.pi * (m - 1) / (L * 2.0 * s)
    return dict(mesh=(m, m, m), xi=xi, rc=s / xi, errortol=errortol,
                P=int(math.ceil(-2.0 * math.log(errortol) / math.pi)))


def _batch_box(args) -> tuple[SystemDefinition, dict]:
    b, n_bodies, body_sizes, phi, seed, errortol = args
    for k in range(8):
        # dense packings of large clusters can cage the growth: take the next seed
        try:
            sysdef = make_system(n_bodies, list(body_sizes), phi, seed=seed + k, charges="uniform", dense_growth=True)
            break
        except RuntimeError:
            continue
    else:
        raise RuntimeError(f"box {b}: cluster growth stalled for 8 seeds at volume fraction {phi:.3f}")
    p = batch_box_params(float(sysdef.L[0]), errortol)
    p["phi"] = phi
    return sysdef, p


def make_batch(n_boxes: int = 64, seed: int = 4, n_range=(2000, 50000), body_sizes=(8, 20, 40),
               phi_range=(0.1, 0.35), errortol: float = 1e-6, workers: int = 1) -> list[tuple[SystemDefinition, dict]]:
    """Seeded set of independent periodic boxes of uneven size: N log-uniform in
    n_range, body sizes drawn per body from body_sizes, volume fraction uniform in
    phi_range, net charges U[-2, 2] shifted to a neutral box; Ewald parameters per box
    from batch_box_params.  Returns [(system, params)] per box, the same for any number
    of worker processes.  The boxes grow with make_system's dense_growth, which reaches
    volume fraction 0.35 with mixed 8/20/40-bead bodies."""
    rng = np.random.default_rng(seed)
    mean_size = float(np.mean(body_sizes))
    jobs = []
    for b in range(n_boxes):
        n_target = math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))
        phi = float(rng.uniform(*phi_range))
        n_bodies = max(2, int(round(n_target / mean_size)))
        jobs.append((b, n_bodies, tuple(body_sizes), phi, int(rng.integers(1 << 30)), errortol))
    if workers <= 1:
        return [_batch_box(j) for j in jobs]
    from concurrent.futures import ProcessPoolExecutor

    with ProcessPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(_batch_box, jobs))
