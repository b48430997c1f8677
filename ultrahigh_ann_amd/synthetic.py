"""Seeded synthetic log-expression data for the benchmark workloads.

Generator (one torch.Generator per call, so a seed fixes the data on a given
device type):
  base level b_j ~ LogNormal(0, 1.5) for each coordinate j < d;
  for each of `classes` classes, a_kj = b_j * exp(0.5 z), z ~ N(0,1), zeroed
  with probability 0.3;
  a sample of class k is a_kj * exp(0.7 N(0,1)), zeroed with probability 0.1,
  mapped to log2(x + 1), stored as binary32.
Values are non-negative and about 63 % of them are nonzero. References and
queries are independent samples with classes drawn uniformly.
"""

from __future__ import annotations

import torch


def class_profiles(d: int, classes: int, gen: torch.Generator, device) -> torch.Tensor:
    base = torch.exp(1.5 * torch.randn(d, generator=gen, device=device))
    prof = base * torch.exp(0.5 * torch.randn(classes, d, generator=gen, device=device))
    keep = torch.rand(classes, d, generator=gen, device=device) >= 0.3
    return prof * keep


def samples(profiles: torch.Tensor, n: int, gen: torch.Generator, chunk: int = 1024) -> torch.Tensor:
    classes, d = profiles.shape
    device = profiles.device
    out = torch.empty(n, d, dtype=torch.float32, device=device)
    labels = torch.randint(classes, (n,), generator=gen, device=device)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        x = profiles[labels[s:e]] * torch.exp(0.7 * torch.randn(e - s, d, generator=gen, device=device))
        x = x * (torch.rand(e - s, d, generator=gen, device=device) >= 0.1)
        out[s:e] = torch.log2(x + 1.0)
    return out


def log_expression(r: int, b: int, d: int = 60660, seed: int = 2026, device="cuda",
                   classes: int = 33):
    """(references r x d, queries b x d) binary32 tensors on `device`."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    prof = class_profiles(d, classes, gen, device)
    refs = samples(prof, r, gen)
    queries = samples(prof, b, gen)
    return refs, queries


def centroid_workload(d: int = 60660, classes: int = 33, per_class: int = 90, queries: int = 3013,
                      seed: int = 2026, device="cuda"):
    """Class-centroid references and held-out queries: for each of `classes`
    classes, `per_class` samples from the generator above are averaged into
    one reference (binary32 mean); the queries are independent samples with
    classes drawn uniformly. Returns (centroids classes x d, queries, query
    classes)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    prof = class_profiles(d, classes, gen, device)
    cents = torch.empty(classes, d, dtype=torch.float32, device=device)
    for k in range(classes):
        x = prof[k].expand(per_class, d) * torch.exp(0.7 * torch.randn(per_class, d, generator=gen, device=device))
        x = x * (torch.rand(per_class, d, generator=gen, device=device) >= 0.1)
        cents[k] = torch.log2(x + 1.0).mean(0)
    labels = torch.randint(classes, (queries,), generator=gen, device=device)
    q = torch.empty(queries, d, dtype=torch.float32, device=device)
    for s0 in range(0, queries, 1024):
        e = min(queries, s0 + 1024)
        x = prof[labels[s0:e]] * torch.exp(0.7 * torch.randn(e - s0, d, generator=gen, device=device))
        x = x * (torch.rand(e - s0, d, generator=gen, device=device) >= 0.1)
        q[s0:e] = torch.log2(x + 1.0)
    return cents, q, labels
