"""Host-driven MC3 sweep with a per-(chain, locus) likelihood callback.

This is the reference's control structure (qupdate, ima_main_mpi.cpp:
1788-2047): chains are visited one by one, each locus is proposed, scored
by one likelihood call and accepted or rejected before the next
(update_gtree.cpp:723-966); alternate sweeps rescale every genealogy of a
chain and re-score all its loci; a swap closes the sweep
(swapchains.cpp:527-662).  The proposals are the same ones the device
sweep uses (kern_mc3.hip), drawn here from numpy.

``lik(c, l, g, dirty, commit)`` returns the log-likelihood of genealogy g
of chain c at locus l; ``dirty`` lists the edges whose partials changed
(-1 for all), ``commit(c, l, accept)`` is called after every decision.
The plain per-locus GPU path and the CPU baseline both run through here.
"""
from __future__ import annotations

import math

import numpy as np


def slide_proposal(rng, g):
    n = g.n
    e = n + int(rng.random() * (n - 1))
    e = min(e, 2 * n - 2)
    a, b = int(g.up0[e]), int(g.up1[e])
    T = g.time[a]
    la = g.time[g.up0[a]] if a >= n else 0.0
    lb = g.time[g.up0[b]] if b >= n else 0.0
    lo = max(la, lb)
    root = g.down[e] == -1
    hi = math.inf if root else g.time[e]
    sigma = 0.5 * max(lo, 0.05) if root else 0.5 * (hi - lo)
    Tn = T + sigma * rng.standard_normal()
    if root:
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


class HostMC3:
    def __init__(self, chains, beta, lik, commit, seed=1, scale_width=0.2, swaptries=None):
        self.chains = chains          # chains[c][l] Genealogy (modified in place)
        self.beta = np.array(beta, float)
        self.lik = lik
        self.commit = commit
        self.rng = np.random.default_rng(seed)
        self.C = len(chains)
        self.L = len(chains[0])
        self.scale_width = scale_width
        self.swaptries = max(1, self.C // 10) if swaptries is None else swaptries
        self.pdg = np.array([[lik(c, l, chains[c][l], (-1,), True) for l in range(self.L)]
                             for c in range(self.C)])
        for c in range(self.C):
            for l in range(self.L):
                commit(c, l, True)
        self.sweeps = 0
        self.stats = np.zeros(6, np.int64)

    def sweep(self, with_scale):
        rng = self.rng
        for c in range(self.C):
            for l in range(self.L):
                g = self.chains[c][l]
                e, a, b, T, Tn = slide_proposal(rng, g)
                g.time[a] = g.time[b] = Tn
                newp = self.lik(c, l, g, (e,), False)
                acc = rng.random() < min(1.0, math.exp(min(700.0, self.beta[c] * (newp - self.pdg[c, l]))))
                if acc:
                    self.pdg[c, l] = newp
                    self.stats[1] += 1
                else:
                    g.time[a] = g.time[b] = T
                self.commit(c, l, acc)
                self.stats[0] += 1
            if with_scale:
                ls = self.scale_width * (rng.random() - 0.5)
                s = math.exp(ls)
                old = [g.time.copy() for g in self.chains[c]]
                newp = np.zeros(self.L)
                a = 0.0
                for l, g in enumerate(self.chains[c]):
                    mask = g.down != -1
                    g.time[mask] = g.time[mask] * s
                    newp[l] = self.lik(c, l, g, (-1,), False)
                    a += self.beta[c] * (newp[l] - self.pdg[c, l]) + (g.n - 1) * ls
                acc = rng.random() < min(1.0, math.exp(min(700.0, a)))
                for l, g in enumerate(self.chains[c]):
                    if acc:
                        self.pdg[c, l] = newp[l]
                    else:
                        g.time[:] = old[l]
                    self.commit(c, l, acc)
                self.stats[2] += 1
                self.stats[3] += int(acc)
        self.swap()
        self.sweeps += 1

    def swap(self):
        C = self.C
        S = self.pdg.sum(axis=1)
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
