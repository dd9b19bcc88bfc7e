"""Seeded synthetic workloads for the five scoring configurations.

Region-level draws (lengths, generating model, tree scale) use numpy's
PCG64 with the configuration's seed; the columns themselves come from the
library's Philox simulator (pcsf_simulate), keyed by the same seed.

  C1  12flies, 1 region x 100 codons, coding model, no gaps; fixed
  C2  58mammals, 10,000 regions U{50..2000}, half coding / half noncoding
      (shuffled), gaps i.i.d. 0.1, 1% N codons; fixed
  C3  100vertebrates, 2,000 regions clip(lognormal(ln 300, 0.8), 30, 5000),
      coding, per-region tree scale U(0.3, 3.0), no gaps; mle
  C4  120mammals, 100,000 regions U{20..600}, 30/70 coding/noncoding,
      gap runs (mean 5 codons) covering 0.25; 6 frames
  C5  100vertebrates, 20,000 regions U{100..3000}, 50/50, gaps 0.15; mle
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .api import Context, PhyloCSFModel


@dataclass
class Workload:
    name: str
    paramset: str
    codes: np.ndarray      # [n_cols, nl] uint8
    offsets: np.ndarray    # [R+1] int64
    region_model: np.ndarray
    region_rho: np.ndarray
    strategy: str

    @property
    def n_cols(self):
        return int(self.offsets[-1])

    @property
    def n_regions(self):
        return len(self.offsets) - 1


SPECS = {
    "C1": dict(paramset="12flies", seed=1, strategy="fixed"),
    "C2": dict(paramset="58mammals", seed=2, strategy="fixed"),
    "C3": dict(paramset="100vertebrates", seed=3, strategy="mle"),
    "C4": dict(paramset="120mammals", seed=4, strategy="fixed"),
    "C5": dict(paramset="100vertebrates", seed=5, strategy="mle"),
}


def region_params(name, n_regions=None):
    """-> lengths, region_model, region_rho, sim kwargs for configuration name."""
    spec = SPECS[name]
    rng = np.random.default_rng(spec["seed"])
    if name == "C1":
        R = 1 if n_regions is None else n_regions
        return np.full(R, 100), np.zeros(R, np.int32), np.ones(R), dict()
    if name == "C2":
        R = 10000 if n_regions is None else n_regions
        lengths = rng.integers(50, 2001, R)
        model = np.zeros(R, np.int32)
        model[R // 2:] = 1
        rng.shuffle(model)
        return lengths, model, np.ones(R), dict(gap_mode=0, gap_p=0.1, n_p=0.01)
    if name == "C3":
        R = 2000 if n_regions is None else n_regions
        lengths = np.clip(np.round(np.exp(rng.normal(np.log(300), 0.8, R))), 30, 5000)
        return (lengths.astype(np.int64), np.zeros(R, np.int32), rng.uniform(0.3, 3.0, R),
                dict())
    if name == "C4":
        R = 100000 if n_regions is None else n_regions
        lengths = rng.integers(20, 601, R)
        model = (rng.random(R) >= 0.3).astype(np.int32)
        return lengths, model, np.ones(R), dict(gap_mode=1, gap_p=0.25, gap_run=5.0)
    if name == "C5":
        R = 20000 if n_regions is None else n_regions
        lengths = rng.integers(100, 3001, R)
        model = np.zeros(R, np.int32)
        model[R // 2:] = 1
        rng.shuffle(model)
        return lengths, model, np.ones(R), dict(gap_mode=0, gap_p=0.15)
    raise KeyError(name)


def make(name, model: PhyloCSFModel, n_regions=None, n_threads=0) -> Workload:
    spec = SPECS[name]
    lengths, rmodel, rrho, kw = region_params(name, n_regions)
    codes, off = model.simulate(lengths, rmodel, rrho, seed=spec["seed"], n_threads=n_threads,
                                **kw)
    return Workload(name, spec["paramset"], codes, off, rmodel, rrho, spec["strategy"])


def model_for(name, ctx: Context) -> PhyloCSFModel:
    return PhyloCSFModel.from_paramset(ctx, SPECS[name]["paramset"])
