"""Helpers shared by the GPU tests of the molecule workloads (the 200-atom
screened molecule and the 100-atom three-centre molecule): a class-stratified
quartet sample, device-side extraction of its output blocks, the oracle
comparison, and the chunked full-output comparison of two device results."""
import os

import numpy as np

from intception_amd import basis
from intception_amd.codegen import canonical_np


def stratified_sample(batch, n_total, seed):
    """About n_total quartet indices, the same number from every canonical
    class present (all of a class when it has fewer), sorted."""
    canon = canonical_np(batch.classes())
    key = canon @ np.array([1000, 100, 10, 1])
    classes = np.unique(key)
    per = max(1, n_total // len(classes))
    rng = np.random.default_rng(seed)
    idx = []
    for k in classes:
        members = np.nonzero(key == k)[0]
        idx.append(members if len(members) <= per else rng.choice(members, size=per, replace=False))
    return np.sort(np.concatenate(idx)), len(classes)


def gather_blocks(out, batch, idx):
    """The output blocks of quartets idx (in that order) from the device
    tensor `out`, as one host array."""
    import torch
    sz = batch.sizes()[idx].astype(np.int64)
    off = batch.out_off[idx].astype(np.int64)
    start = np.repeat(off - np.concatenate([[0], np.cumsum(sz)[:-1]]), sz)
    el = np.arange(int(sz.sum()), dtype=np.int64) + start
    return out[torch.from_numpy(el).to(out.device)].cpu().numpy()


def oracle_blocks(batch, idx, threads=16):
    """Oracle values (accurate Boys scheme) of quartets idx, packed in order."""
    from oracle import oracle
    sub = batch.subset(idx)
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    return oracle.eri_batch(sub, True, basis.SPH, min(threads, os.cpu_count() or 1)), sub


def block_bar_check(got, ref, sub, rtol=1e-12, atol=1e-14, block_rel=0.0):
    """Per quartet: |got - ref| <= atol + rtol |ref| + block_rel max|ref block|.
    Returns (number of elements outside, worst ratio err / bar, max abs err)."""
    sz = sub.sizes().astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sz)[:-1]])
    bmax = np.maximum.reduceat(np.abs(ref), off) if len(ref) else np.zeros(0)
    bar = atol + rtol * np.abs(ref) + block_rel * np.repeat(bmax, sz)
    err = np.abs(got - ref)
    return int(np.sum(~(err <= bar))), float(np.max(err / bar)) if len(err) else 0.0, float(err.max(initial=0.0))


def max_abs_diff(a, b, chunk=1 << 28):
    """max |a - b| over two equal-length device tensors, in chunks."""
    m = 0.0
    for s0 in range(0, a.numel(), chunk):
        m = max(m, float((a[s0:s0 + chunk] - b[s0:s0 + chunk]).abs().max().item()))
    return m
