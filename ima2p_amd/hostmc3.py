"""Host-driven MC3 sweep with per-(chain, locus) likelihood and prior callbacks.

This is the reference's control structure (qupdate, ima_main_mpi.cpp:
1788-2047): chains are visited one by one, each locus is proposed, scored
and accepted or rejected before the next (update_gtree.cpp:723-966);
alternate sweeps rescale every genealogy of a chain -- globally, or with
``ry1`` by changet_RY1's rubber band on one split time (update_t_RY.cpp:
223-517) -- and re-score all its loci; a swap closes the sweep
(swapchains.cpp:527-662).  The proposals are
the ones the device sweep uses (kern_mc3.hip), drawn here from numpy.

``lik(c, l, g, dirty, commit)`` returns the log-likelihood of genealogy g
of chain c at locus l; ``dirty`` lists the edges whose partials changed
(-1 for all), ``commit(c, l, accept)`` is called after every decision.
With a ``prior`` (see ``HostPrior``) acceptance is beta (dprobg + dpdg)
and the proposals respect split times and migrations.  The plain
per-locus GPU path and the CPU baseline both run through here.
"""
from __future__ import annotations

import math

import numpy as np


def slide_bounds(g, e, tv=None):
    """Interval a node may move in: its daughters' and parent's times, and
    with a population model (tv = split times) its period and the
    migrations on its daughter edges (below) and its parent edge (above)."""
    n = g.n
    a, b = int(g.up0[e]), int(g.up1[e])
    T = g.time[a]
    lo = max(g.time[g.up0[a]] if a >= n else 0.0, g.time[g.up0[b]] if b >= n else 0.0)
    hi = math.inf if g.down[e] == -1 else g.time[e]
    if tv is not None:
        k = 0
        while k < len(tv) and tv[k] <= T:
            k += 1
        if k > 0:
            lo = max(lo, tv[k - 1])
        if k < len(tv):
            hi = min(hi, tv[k])
        for t, _ in g.mig[a]:
            lo = max(lo, t)
        for t, _ in g.mig[b]:
            lo = max(lo, t)
        for t, _ in g.mig[e]:
            hi = min(hi, t)
    return a, b, T, lo, hi


def slide_proposal(rng, g, tv=None):
    n = g.n
    e = n + int(rng.random() * (n - 1))
    e = min(e, 2 * n - 2)
    a, b, T, lo, hi = slide_bounds(g, e, tv)
    open_ = hi == math.inf
    sigma = 0.5 * max(lo, 0.05) if open_ else 0.5 * (hi - lo)
    Tn = T + sigma * rng.standard_normal()
    if open_:
        Tn = lo + abs(Tn - lo)
    else:
        w = hi - lo
        y = math.fmod(Tn - lo, 2 * w)
        if y < 0:
            y += 2 * w
        if y > w:
            y = 2 * w - y
        Tn = lo + y
    if not (lo < Tn < hi):
        Tn = T
    return e, a, b, T, Tn


class HostPrior:
    """Interface of the genealogy prior for HostMC3: all-loci state per chain.

    init(c, gens, tv) -> probg; try_locus(c, l, g) -> probg with locus l
    replaced by g; settle_locus(c, l, accept); try_chain(c, gens, tv) ->
    probg with every locus and the split times replaced; settle_chain(c,
    accept)."""


class EnginePrior(HostPrior):
    """The prior through the engine's batch entry point, one call per
    proposal: the chain's genealogies live in slots[c][l]."""

    def __init__(self, engine, model, slots):
        self.e = engine
        self.model = model
        self.slots = slots
        self.tv = {}

    def _eval(self, c, tv):
        pg, _, _ = self.e.prior([self.slots[c]], np.asarray(tv, float)[None, :], self.model)
        return float(pg[0])

    def init(self, c, gens, tv):
        for l, g in enumerate(gens):
            self.e.write_slots(self.slots[c][l], [g])
            self.e.write_pops(self.slots[c][l], [g])
        self.tv[c] = np.array(tv, float)
        return self._eval(c, tv)

    def try_locus(self, c, l, g):
        self.e.write_slots(self.slots[c][l], [g])
        return self._eval(c, self.tv[c])

    def settle_locus(self, c, l, accept, old=None):
        if not accept and old is not None:
            self.e.write_slots(self.slots[c][l], [old])

    def try_chain(self, c, gens, tv):
        for l, g in enumerate(gens):
            self.e.write_slots(self.slots[c][l], [g])
            self.e.write_pops(self.slots[c][l], [g])
        return self._eval(c, tv)

    def settle_chain(self, c, accept, old=None, tv=None):
        if accept:
            self.tv[c] = np.array(tv, float)
        elif old is not None:
            for l, g in enumerate(old):
                self.e.write_slots(self.slots[c][l], [g])
                self.e.write_pops(self.slots[c][l], [g])


class HostMC3:
    def __init__(self, chains, beta, lik, commit, seed=1, scale_width=0.2, swaptries=None,
                 prior=None, tv=None, tmax=math.inf, ry1=False):
        self.chains = chains          # chains[c][l] Genealogy (modified in place)
        self.beta = np.array(beta, float)
        self.lik = lik
        self.commit = commit
        self.rng = np.random.default_rng(seed)
        self.C = len(chains)
        self.L = len(chains[0])
        self.scale_width = scale_width
        self.ry1 = ry1
        self.swaptries = max(1, self.C // 10) if swaptries is None else swaptries
        self.prior = prior
        self.tmax = tmax
        self.tv = [None if tv is None else np.array(tv, float) for _ in range(self.C)]
        self.pdg = np.array([[lik(c, l, chains[c][l], (-1,), True) for l in range(self.L)]
                             for c in range(self.C)])
        for c in range(self.C):
            for l in range(self.L):
                commit(c, l, True)
        self.probg = np.zeros(self.C)
        if prior is not None:
            for c in range(self.C):
                self.probg[c] = prior.init(c, chains[c], self.tv[c])
        self.sweeps = 0
        self.stats = np.zeros(6, np.int64)

    def _locus(self, c, l):
        rng = self.rng
        g = self.chains[c][l]
        tv = self.tv[c] if self.prior is not None else None
        e, a, b, T, Tn = slide_proposal(rng, g, tv)
        g.time[a] = g.time[b] = Tn
        newp = self.lik(c, l, g, (e,), False)
        dprior = 0.0
        if self.prior is not None:
            pn = self.prior.try_locus(c, l, g)
            dprior = pn - self.probg[c]
        x = self.beta[c] * (dprior + newp - self.pdg[c, l])
        acc = rng.random() < min(1.0, math.exp(min(700.0, x)))
        if acc:
            self.pdg[c, l] = newp
            if self.prior is not None:
                self.probg[c] = pn
            self.stats[1] += 1
        else:
            g.time[a] = g.time[b] = T
        if self.prior is not None:
            self.prior.settle_locus(c, l, acc, None if acc else g)
        self.commit(c, l, acc)
        self.stats[0] += 1

    def _rescale(self, c):
        rng = self.rng
        ls = self.scale_width * (rng.random() - 0.5)
        s = math.exp(ls)
        gens = self.chains[c]
        old = [g.copy() for g in gens]
        withp = self.prior is not None
        tv_new = self.tv[c] * s if withp and self.tv[c] is not None else self.tv[c]
        newp = np.zeros(self.L)
        a = 0.0
        for l, g in enumerate(gens):
            mask = g.down != -1
            g.time[mask] = g.time[mask] * s
            nscaled = g.n - 1
            if withp:
                g.mig = [[(t * s, q) for t, q in m] for m in g.mig]
                nscaled += sum(len(m) for m in g.mig)
            newp[l] = self.lik(c, l, g, (-1,), False)
            a += self.beta[c] * (newp[l] - self.pdg[c, l]) + nscaled * ls
        ok = True
        if withp:
            pn = self.prior.try_chain(c, gens, tv_new)
            a += self.beta[c] * (pn - self.probg[c]) + len(tv_new) * ls
            ok = bool(np.all(tv_new <= self.tmax))
        acc = ok and rng.random() < min(1.0, math.exp(min(700.0, a)))
        for l, g in enumerate(gens):
            if acc:
                self.pdg[c, l] = newp[l]
            else:
                g.time[:] = old[l].time
                g.mig = old[l].mig
            self.commit(c, l, acc)
        if withp:
            if acc:
                self.probg[c] = pn
                self.tv[c] = tv_new
            self.prior.settle_chain(c, acc, None if acc else gens, tv_new)
        self.stats[2] += 1
        self.stats[3] += int(acc)

    def _rubber_band(self, c):
        """changet_RY1 (update_t_RY.cpp:223-517): split k uniform, the new
        time from getnewt's reflected window (update_gtree_common.cpp:
        2501-2518), times between the neighbouring splits stretched by
        beforesplit / aftersplit (update_t_RY.cpp:59-82), Hastings term
        (ecd + emd) log t_d_hterm + (ecu + emu) log t_u_hterm."""
        rng = self.rng
        gens = self.chains[c]
        tv = self.tv[c]
        ns = len(tv)
        k = min(int(rng.random() * ns), ns - 1)
        oldt = float(tv[k])
        t_u = 0.0 if k == 0 else float(tv[k - 1])
        t_d = 1e6 if k == ns - 1 else float(tv[k + 1])
        tdp, tup = min(self.tmax, t_d), max(0.0, t_u)
        twin = (tdp - tup) / (math.log(self.L + 1.0) * (ns + 1 - k))
        newt = (oldt - twin / 2) + rng.random() * twin
        if newt >= tdp:
            newt = 2.0 * tdp - newt
        elif newt <= tup:
            newt = 2.0 * tup - newt

        def move(t):
            if t_u < t <= oldt:
                return (t * newt / oldt if k == 0 else
                        t_u + (t - t_u) * (newt - t_u) / (oldt - t_u)), 1, 0
            if oldt < t < t_d:
                return (t + newt - oldt if k == ns - 1 else
                        t_d - (t_d - newt) * (t_d - t) / (t_d - oldt)), 0, 1
            return t, 0, 0

        old = [g.copy() for g in gens]
        ecu = ecd = emu = emd = 0
        newp = np.zeros(self.L)
        a = 0.0
        for l, g in enumerate(gens):
            for i in range(2 * g.n - 1):
                if g.down[i] != -1:
                    g.time[i], u, d = move(float(g.time[i]))
                    ecu += u
                    ecd += d
            mig = []
            for m in g.mig:
                row = []
                for t, q in m:
                    t2, u, d = move(t)
                    emu += u
                    emd += d
                    row.append((t2, q))
                mig.append(row)
            g.mig = mig
            newp[l] = self.lik(c, l, g, (-1,), False)
            a += self.beta[c] * (newp[l] - self.pdg[c, l])
        ecu //= 2
        ecd //= 2
        tv_new = tv.copy()
        tv_new[k] = newt
        pn = self.prior.try_chain(c, gens, tv_new)
        t_u_hterm = (newt - t_u) / (oldt - t_u)
        t_d_hterm = 1.0 if k == ns - 1 else (t_d - newt) / (t_d - oldt)
        a += (self.beta[c] * (pn - self.probg[c]) + (ecd + emd) * math.log(t_d_hterm)
              + (ecu + emu) * math.log(t_u_hterm))
        acc = math.log(max(rng.random(), 1e-300)) < min(1.0, a)
        for l, g in enumerate(gens):
            if acc:
                self.pdg[c, l] = newp[l]
            else:
                g.time[:] = old[l].time
                g.mig = old[l].mig
            self.commit(c, l, acc)
        if acc:
            self.probg[c] = pn
            self.tv[c] = tv_new
        self.prior.settle_chain(c, acc, None if acc else gens, tv_new)
        self.stats[2] += 1
        self.stats[3] += int(acc)

    def sweep(self, with_scale):
        for c in range(self.C):
            for l in range(self.L):
                self._locus(c, l)
            if with_scale:
                if self.ry1 and self.prior is not None and self.tv[c] is not None \
                        and len(self.tv[c]) > 0:
                    self._rubber_band(c)
                else:
                    self._rescale(c)
        self.swap()
        self.sweeps += 1

    def swap(self):
        C = self.C
        S = self.pdg.sum(axis=1) + self.probg
        for _ in range(self.swaptries if C > 1 else 0):
            ci = int(self.rng.random() * C)
            if C < 17:
                cjmin, cjrange = 0, C
            else:
                cjmin = max(0, ci - 7)
                cjrange = min(C, ci + 7) - cjmin
            cj = ci
            while cj == ci:
                cj = cjmin + int(self.rng.random() * cjrange)
            w = math.exp(min(700.0, (self.beta[ci] - self.beta[cj]) * (S[cj] - S[ci])))
            self.stats[4] += 1
            if w >= 1.0 or w > self.rng.random():
                self.beta[ci], self.beta[cj] = self.beta[cj], self.beta[ci]
                self.stats[5] += 1
