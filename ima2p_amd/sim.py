"""Synthetic Isolation-with-Migration data in the reference's own layouts.

Genealogies are simulated backward in time under the structured coalescent
with the reference's scaling: time is mutation-scaled, a pair of lineages in
population p coalesces at rate 2/(h * theta_p) (update_gtree_common.cpp:1853),
a lineage in p moves to q (coalescent direction) at rate m_{p>q}
(update_gtree_common.cpp:1858-1862), and populations merge at the split
times of the population tree.

Each genealogy is returned in the reference's edge layout (imamp.hpp:651-679):
edges 0..n-1 are the sampled gene copies, n..2n-2 internal; ``up0/up1`` are
daughter edges, ``down`` the parent edge (-1 at the root edge), ``time`` the
time at the bottom of the edge (TIMEMAX on the root edge), ``pop`` the
population at the top of the edge and ``mig`` the migration events on it.

Data encodings follow readata.cpp:
  * HKY: empirical base frequencies over all read bases, gap columns dropped,
    identical columns compressed in first-occurrence order with counts
    (``readseqHKY`` readata.cpp:246-344, ``sortseq`` :202-221);
  * IS: variable biallelic sites coded 0/1 relative to the first gene copy
    (``findsegsites`` :34-177, ``readseqIS`` :347-495);
  * SW: integer allele sizes per gene copy (``readseqSW`` :500-588).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

TIMEMAX = 1000000.0  # imamp.hpp:129


@dataclass
class PopTree:
    """A population tree: sampled pops 0..npops-1, ancestral npops..2npops-2.

    ``joins[k] = (a, b)`` merges a and b into population npops + k at split
    time index k (the reference's ``droppops``/``addpop``, build_poptree.cpp).
    """
    npops: int
    joins: list

    @staticmethod
    def two_pop() -> "PopTree":
        return PopTree(2, [(0, 1)])

    @staticmethod
    def three_pop_01_2() -> "PopTree":
        # ((0,1):3,2):4
        return PopTree(3, [(0, 1), (3, 2)])

    @property
    def ntreepops(self) -> int:
        return 2 * self.npops - 1

    def active(self, period: int) -> list:
        pops = set(range(self.npops))
        for k in range(period):
            a, b = self.joins[k]
            pops.discard(a)
            pops.discard(b)
            pops.add(self.npops + k)
        return sorted(pops)


@dataclass
class Genealogy:
    n: int
    up0: np.ndarray
    up1: np.ndarray
    down: np.ndarray
    time: np.ndarray
    root: int
    pop: np.ndarray
    mig: list = field(default_factory=list)  # per edge: list of (mt, mp)

    def branch_lengths(self) -> np.ndarray:
        bl = np.zeros(2 * self.n - 1)
        for e in range(2 * self.n - 1):
            if self.down[e] == -1:
                continue
            u = self.up0[e]
            bl[e] = self.time[e] - (0.0 if u == -1 else self.time[u])
        return bl

    def length(self) -> float:
        return float(self.branch_lengths().sum())

    def copy(self) -> "Genealogy":
        return Genealogy(self.n, self.up0.copy(), self.up1.copy(),
                         self.down.copy(), self.time.copy(), int(self.root),
                         self.pop.copy(), [list(m) for m in self.mig])

    def clades(self) -> list:
        """Python int bitmask of tips below each edge."""
        m = [0] * (2 * self.n - 1)
        for t in range(self.n):
            e = t
            while e != -1:
                m[e] |= 1 << t
                e = int(self.down[e])
        return m

    def check(self) -> None:
        n = self.n
        assert self.down[self.root] == -1
        for e in range(2 * n - 1):
            if e < n:
                assert self.up0[e] == -1 and self.up1[e] == -1
            else:
                a, b = self.up0[e], self.up1[e]
                assert self.down[a] == e and self.down[b] == e
                assert self.time[a] == self.time[b]
                if self.down[e] != -1:
                    assert self.time[e] >= self.time[a]


def simulate_genealogy(rng: np.random.Generator, samppop, theta, tsplit,
                       mig, poptree: PopTree, h: float = 1.0) -> Genealogy:
    """Structured coalescent with splits.

    samppop: gene copies per sampled population; theta: per tree population
    (length 2npops-1); tsplit: split times (length npops-1, increasing);
    mig: (2npops-1)x(2npops-1) coalescent-direction migration rates.
    """
    samppop = list(samppop)
    n = int(sum(samppop))
    N = 2 * n - 1
    up0 = -np.ones(N, np.int32)
    up1 = -np.ones(N, np.int32)
    down = -np.ones(N, np.int32)
    time = np.zeros(N)
    pop = np.zeros(N, np.int32)
    migs = [[] for _ in range(N)]
    lin_edge, lin_pop = [], []
    e = 0
    for p, cnt in enumerate(samppop):
        for _ in range(cnt):
            lin_edge.append(e)
            lin_pop.append(p)
            pop[e] = p
            e += 1
    nxt = n
    t = 0.0
    period = 0
    nsplit = poptree.npops - 1
    mig = np.asarray(mig, float)
    while len(lin_edge) > 1:
        active = poptree.active(period)
        bound = tsplit[period] if period < nsplit else np.inf
        counts = {p: 0 for p in active}
        for p in lin_pop:
            counts[p] += 1
        rates = []
        for p in active:
            k = counts[p]
            rates.append(("c", p, None, k * (k - 1) / (h * theta[p])))
            for q in active:
                if q != p and mig[p, q] > 0 and k > 0:
                    rates.append(("m", p, q, k * mig[p, q]))
        tot = sum(r[3] for r in rates)
        dt = rng.exponential(1.0 / tot) if tot > 0 else np.inf
        if t + dt >= bound:
            t = bound
            a, b = poptree.joins[period]
            newp = poptree.npops + period
            lin_pop = [newp if p in (a, b) else p for p in lin_pop]
            period += 1
            continue
        t += dt
        w = np.array([r[3] for r in rates])
        kind, p, q, _ = rates[rng.choice(len(rates), p=w / w.sum())]
        idx = [i for i, pp in enumerate(lin_pop) if pp == p]
        if kind == "c":
            i, j = sorted(rng.choice(idx, 2, replace=False))
            a, b = lin_edge[i], lin_edge[j]
            time[a] = time[b] = t
            down[a] = down[b] = nxt
            up0[nxt], up1[nxt] = a, b
            pop[nxt] = p
            del lin_edge[j], lin_pop[j]
            lin_edge[i] = nxt
            lin_pop[i] = p
            nxt += 1
        else:
            i = int(rng.choice(idx))
            migs[lin_edge[i]].append((t, int(q)))
            lin_pop[i] = q
    root = lin_edge[0]
    time[root] = TIMEMAX
    g = Genealogy(n, up0, up1, down, time, root, pop, migs)
    return g


def relabel_internal(g: Genealogy, rng: np.random.Generator) -> Genealogy:
    """Renumber internal edges at random (the reference does not keep them
    in time order after genealogy updates)."""
    n = g.n
    perm = np.arange(2 * n - 1)
    perm[n:] = n + rng.permutation(n - 1)  # old -> new
    N = 2 * n - 1
    up0 = -np.ones(N, np.int32)
    up1 = -np.ones(N, np.int32)
    down = -np.ones(N, np.int32)
    time = np.zeros(N)
    pop = np.zeros(N, np.int32)
    mig = [None] * N
    for old in range(N):
        new = perm[old]
        up0[new] = perm[g.up0[old]] if g.up0[old] >= 0 else -1
        up1[new] = perm[g.up1[old]] if g.up1[old] >= 0 else -1
        down[new] = perm[g.down[old]] if g.down[old] >= 0 else -1
        time[new] = g.time[old]
        pop[new] = g.pop[old]
        mig[new] = list(g.mig[old])
    return Genealogy(n, up0, up1, down, time, int(perm[g.root]), pop, mig)


# ------------------------------------------------------ genealogy moves --

def move_edge(g: Genealogy, edge: int, newsis: int, newtime: float) -> Genealogy:
    """The reference's genealogy move (update_gtree.cpp:760-800): detach
    `edge` with its parent node, join the old sister to the parent's parent
    (joinsisdown), and re-attach `edge` on `newsis` at `newtime` through the
    freed node (splitsisdown).  `time` is at the rootward end of an edge;
    newtime must lie above both edges' upper ends and below newsis' lower
    end.  Population labels and migrations are left as they were (single
    population use)."""
    h = g.copy()
    down = h.down[edge]
    assert down != -1 and newsis not in (edge, down)
    sis = h.up1[down] if h.up0[down] == edge else h.up0[down]
    par = h.down[down]
    # joinsisdown
    h.down[sis] = par
    h.time[sis] = h.time[down]
    if par == -1:
        h.root = int(sis)
    elif h.up0[par] == down:
        h.up0[par] = sis
    else:
        h.up1[par] = sis
    # splitsisdown
    upper = lambda e: h.time[h.up0[e]] if h.up0[e] != -1 else 0.0  # noqa: E731
    assert newtime > upper(edge) and newtime > upper(newsis) and newtime < h.time[newsis]
    q = h.down[newsis]
    h.up0[down], h.up1[down] = edge, newsis
    h.down[down] = q
    h.time[down] = h.time[newsis]
    h.time[newsis] = newtime
    h.time[edge] = newtime
    h.down[edge] = h.down[newsis] = down
    if q == -1:
        h.root = int(down)
    elif h.up0[q] == newsis:
        h.up0[q] = down
    else:
        h.up1[q] = down
    return h


@dataclass
class SWMove:
    """Inputs of finishSWupdateA (update_gtree_common.cpp:2219-2362) for one
    move: the edges, the stored copies (storeAinfo, :820-846) and the draws
    of geometric() and bitran()."""
    edge: int
    downedge: int
    sisedge: int
    newsisedge: int
    copyA0: int
    copyA1: int
    down_was_root: bool
    hold_down_A: int
    minA: int
    maxA: int
    flip: int
    copy_dlikeA: tuple
    newsis_dlikeA: float
    u: float
    uni: float
    slot: int = 0


def random_sw_move(rng, g: Genealogy, A, dlikeA, u, minA, maxA):
    """A random valid move on g with the SWMove that scores it; returns
    (genealogy after the move, move)."""
    N = 2 * g.n - 1
    while True:
        edge = int(rng.integers(0, N))
        if g.down[edge] == -1:
            continue
        down = int(g.down[edge])
        sis = int(g.up1[down] if g.up0[down] == edge else g.up0[down])
        cands = [e for e in range(N) if e not in (edge, down)]
        newsis = int(rng.choice(cands))
        # the tree after joinsisdown decides newsis' span
        t_lo = lambda e: g.time[g.up0[e]] if g.up0[e] != -1 else 0.0  # noqa: E731
        lo = max(t_lo(edge), t_lo(newsis))
        hi = g.time[down] if newsis == sis else g.time[newsis]
        if newsis == g.root or (newsis == sis and g.down[down] == -1):
            hi = lo + 2.0 * (1.0 + lo)
        if not hi > lo:
            continue
        newtime = float(rng.uniform(lo, hi))
        if newtime <= lo or newtime >= hi:
            continue
        h = move_edge(g, edge, newsis, newtime)
        par = int(g.down[down])
        mv = SWMove(edge, down, sis, newsis, int(A[edge]), int(A[sis]), par == -1,
                    int(A[par]) if par != -1 else -1, minA, maxA, int(rng.integers(0, 2)),
                    (float(dlikeA[edge]), float(dlikeA[sis]),
                     float(dlikeA[down]) if par != -1 else 0.0),
                    float(dlikeA[newsis]), float(u), float(rng.uniform(1e-12, 1.0)))
        return h, mv


# ---------------------------------------------------------------- HKY data --

def hky_pmatrix(pi, kappa, mt):
    """HKY transition matrix P[from, to] for mt = mutrate * t (own
    statement of the closed form; class sums of the destination base)."""
    pi = np.asarray(pi, float)
    P = np.zeros((4, 4))
    e = np.exp(-mt)
    for to in range(4):
        PI = pi[0] + pi[2] if to in (0, 2) else pi[1] + pi[3]
        A = 1.0 + PI * (kappa - 1.0)
        eA = np.exp(-mt * A)
        for fr in range(4):
            if fr == to:
                P[fr, to] = pi[to] + pi[to] * e * (1 / PI - 1) + eA * (PI - pi[to]) / PI
            elif (fr + to) in (2, 4):
                P[fr, to] = pi[to] + pi[to] * e * (1 / PI - 1) - pi[to] / PI * eA
            else:
                P[fr, to] = pi[to] * (1 - e)
    return P


def standfactor(pi, kappa):
    s = 0.0
    for i in range(4):
        for j in range(4):
            if i != j:
                s += pi[i] * pi[j] * (kappa if (i + j) in (2, 4) else 1.0)
    return s


def simulate_hky_sequences(rng, g: Genealogy, nsites, pi, kappa, u):
    """Evolve nsites bases down the genealogy; returns [n, nsites] in 0..3."""
    pi = np.asarray(pi, float)
    mutrate = u / (nsites * standfactor(pi, kappa))
    N = 2 * g.n - 1
    states = np.zeros((N, nsites), np.int8)
    states[g.root] = rng.choice(4, nsites, p=pi)
    # top-down: process edges from the root by decreasing bottom time
    order = [g.root]
    out = []
    while order:
        e = order.pop()
        out.append(e)
        if g.up0[e] >= 0:
            order += [int(g.up0[e]), int(g.up1[e])]
    bl = g.branch_lengths()
    for e in out:
        if e == g.root or g.up0[e] < 0:
            pass
        for c in ((int(g.up0[e]), int(g.up1[e])) if g.up0[e] >= 0 else ()):
            # the node at the top of e holds states[e]; child c evolves along
            # its own branch
            P = hky_pmatrix(pi, kappa, mutrate * bl[c])
            cum = np.cumsum(P, axis=1)
            r = rng.random(nsites)
            par = states[e]
            states[c] = (r[:, None] > cum[par]).sum(axis=1).clip(0, 3)
    return states[: g.n].astype(np.int8)


@dataclass
class HKYLocus:
    n: int
    seq: np.ndarray      # [n, S] int8 0..3 (compressed patterns)
    mult: np.ndarray     # [S] int32
    totsites: int
    pi: np.ndarray       # [4] empirical


def encode_hky(raw: np.ndarray) -> HKYLocus:
    """readseqHKY: raw[n, L] with -1 for gaps/N."""
    raw = np.asarray(raw)
    n, L = raw.shape
    counts = np.array([(raw == b).sum() for b in range(4)], float)
    keep = ~(raw < 0).any(axis=0)
    cols = raw[:, keep]
    totsites = cols.shape[1]
    seen = {}
    order = []
    mult = []
    for j in range(totsites):
        key = cols[:, j].tobytes()
        if key in seen:
            mult[seen[key]] += 1
        else:
            seen[key] = len(order)
            order.append(j)
            mult.append(1)
    seq = cols[:, order].astype(np.int8)
    return HKYLocus(n, seq, np.array(mult, np.int32), int(totsites),
                    counts / counts.sum())


# ----------------------------------------------------------------- IS data --

@dataclass
class ISLocus:
    n: int
    seq: np.ndarray      # [n, S] int8 0/1 relative to gene copy 0


def simulate_is(rng, g: Genealogy, u) -> ISLocus:
    bl = g.branch_lengths()
    clades = g.clades()
    sites = []
    for e in range(2 * g.n - 1):
        if g.down[e] == -1:
            continue
        k = rng.poisson(u * bl[e])
        for _ in range(k):
            sites.append(clades[e])
    rng.shuffle(sites)
    S = len(sites)
    seq = np.zeros((g.n, S), np.int8)
    for j, m in enumerate(sites):
        for t in range(g.n):
            seq[t, j] = (m >> t) & 1
        if seq[0, j] == 1:
            seq[:, j] ^= 1
    return ISLocus(g.n, seq)


# ----------------------------------------------------------------- SW data --

def simulate_sw(rng, g: Genealogy, u, start=20):
    """Stepwise allele walk: Poisson(u t) steps of +-1 on every branch.
    Returns allele states for all 2n-1 edges (state at the top of each)."""
    N = 2 * g.n - 1
    A = np.zeros(N, np.int32)
    A[g.root] = start
    bl = g.branch_lengths()
    stack = [g.root]
    while stack:
        e = stack.pop()
        if g.up0[e] < 0:
            continue
        for c in (int(g.up0[e]), int(g.up1[e])):
            k = rng.poisson(u * bl[c])
            A[c] = A[e] + int(rng.choice([-1, 1], k).sum()) if k else A[e]
            stack.append(c)
    return A


# --------------------------------------------------------- model configs --

def im_model_two_pop(theta=(5.0, 5.0, 5.0), t=1.0, m12=0.5, m21=0.5):
    pt = PopTree.two_pop()
    mig = np.zeros((3, 3))
    mig[0, 1] = m12
    mig[1, 0] = m21
    return pt, np.array(theta, float), np.array([t]), mig


def im_model_three_pop(rng):
    """Tree ((0,1),2): theta U(1,10) for 5 populations, split times U(0.5,3)
    sorted, migration U(0,1) between every pair that coexists."""
    pt = PopTree.three_pop_01_2()
    theta = rng.uniform(1, 10, 5)
    ts = np.sort(rng.uniform(0.5, 3, 2))
    mig = np.zeros((5, 5))
    for period in range(2):
        act = pt.active(period)
        for p in act:
            for q in act:
                if p != q:
                    mig[p, q] = rng.uniform(0, 1)
    return pt, theta, ts, mig


def im_model_random(rng, P):
    """P populations on the ladder tree (((0,1),2),...): theta U(1,10) for the
    2P-1 populations, split times U(0.5,3) sorted, migration U(0,1) between
    every pair that coexists before the last split.  P = 3 gives the tree
    of im_model_three_pop."""
    joins = [(0, 1)] + [(P + k - 1, k + 1) for k in range(1, P - 1)]
    pt = PopTree(P, joins)
    theta = rng.uniform(1, 10, 2 * P - 1)
    ts = np.sort(rng.uniform(0.5, 3, P - 1))
    mig = np.zeros((2 * P - 1, 2 * P - 1))
    for period in range(P - 1):
        act = pt.active(period)
        for p in act:
            for q in act:
                if p != q:
                    mig[p, q] = rng.uniform(0, 1)
    return pt, theta, ts, mig


# ------------------------------------------------------- L-mode tables --

def lmode_table(rng, G, npsize=3, nmig=2, nsplit=1, theta=5.0, m=0.5, nloci=5):
    """Genealogy summary rows in the savegsampinf layout (ginfo.cpp:288-316):
    cc | fc | hcc (npsize) | mc | fm (nmig) | qintegrate (npsize) |
    mintegrate (nmig) | pdg | probg | t (nsplit).  Coalescent and migration
    counts are Poisson, interval sums exponential-like, hcc = 0 (h = 1)."""
    K = 4 * npsize + 3 * nmig + 2 + nsplit
    tab = np.zeros((G, K), np.float32)
    cc = rng.poisson(4.0 * nloci, (G, npsize)).astype(np.float64)
    fc = rng.gamma(cc + 1.0, theta / 2.0)
    mc = rng.poisson(2.0 * nloci * m, (G, nmig)).astype(np.float64)
    fm = rng.gamma(mc + 1.0, 1.0 / m)
    o = 0
    tab[:, o:o + npsize] = cc; o += npsize
    tab[:, o:o + npsize] = fc; o += npsize
    tab[:, o:o + npsize] = 0.0; o += npsize
    tab[:, o:o + nmig] = mc; o += nmig
    tab[:, o:o + nmig] = fm; o += nmig
    tab[:, o:o + npsize] = rng.normal(-20.0, 2.0, (G, npsize)); o += npsize
    tab[:, o:o + nmig] = rng.normal(-8.0, 1.0, (G, nmig)); o += nmig
    tab[:, o] = rng.normal(-500.0, 10.0, G); o += 1
    tab[:, o] = tab[:, 4 * npsize + 2 * nmig:4 * npsize + 3 * nmig].sum(1) + \
        tab[:, 3 * npsize + 2 * nmig:4 * npsize + 2 * nmig].sum(1); o += 1
    tab[:, o:o + nsplit] = rng.uniform(0.5, 2.0, (G, nsplit))
    return tab


def lmode_grid(lo, hi, k, dims=5):
    """k^dims grid points over [lo, hi] per dimension (row-major)."""
    axes = [np.linspace(lo[d], hi[d], k) for d in range(dims)]
    mesh = np.meshgrid(*axes, indexing="ij")
    return np.stack([m.ravel() for m in mesh], axis=1)
