"""Ragged mixed-width modexp workloads (config 4) on the device.

A workload is a list of sub-batches, each with its own odd modulus N (64 or
128 limbs), bases below N and per-item exponents, all resident on the GPU in
the limb-interleaved layout. ``RaggedPlan`` (engine.py) runs it as one
persistent hipGraph-captured launch; ``run_plain`` runs one launch per
sub-batch for comparison.
"""

from __future__ import annotations

import math
import random
from dataclasses import dataclass

import numpy as np

from .engine import Context, Modulus, stride_for
from .limbs import limbs_to_ints


@dataclass
class SubBatch:
    modulus: int
    limbs: int
    count: int
    e_bits: int
    bases: np.ndarray   # (count, limbs) uint32, host copy
    exps: np.ndarray    # (count, e_limbs) uint32, host copy


def random_modulus(rng: random.Random, n_bits: int) -> int:
    """N = n^2 for a seeded odd n of n_bits bits (top bit set). Montgomery
    arithmetic only needs N odd; the factors need not be prime."""
    n = rng.getrandbits(n_bits) | 1 | (1 << (n_bits - 1))
    return n * n


def make_subbatches(n_subs: int, max_size: int, seed: int, n_bits_choices=(1024, 2048),
                    min_size: int = 1) -> list[SubBatch]:
    """Sizes log-uniform in [min_size, max_size]; width chosen 50/50 per
    sub-batch; bases uniform below N (limb-wise: top limb below N's top limb);
    exponents of full modulus width (top bit set)."""
    rng = random.Random(seed)
    nrng = np.random.default_rng(seed)
    out = []
    for _ in range(n_subs):
        n_bits = rng.choice(n_bits_choices)
        N = random_modulus(rng, n_bits)
        limbs = 64 if N.bit_length() <= 2048 else 128
        size = int(math.floor(math.exp(rng.uniform(math.log(min_size), math.log(max_size + 1)))))
        size = max(min_size, min(max_size, size))
        top = (N >> (32 * (limbs - 1))) & 0xFFFFFFFF
        bases = nrng.integers(0, 2 ** 32, size=(size, limbs), dtype=np.uint64).astype(np.uint32)
        bases[:, limbs - 1] = (bases[:, limbs - 1].astype(np.uint64) % max(top, 1)).astype(np.uint32)
        e_bits = N.bit_length()
        e_limbs = (e_bits + 31) // 32
        exps = nrng.integers(0, 2 ** 32, size=(size, e_limbs), dtype=np.uint64).astype(np.uint32)
        hi_bits = e_bits - 32 * (e_limbs - 1)
        exps[:, e_limbs - 1] &= np.uint32((1 << hi_bits) - 1) if hi_bits < 32 else np.uint32(0xFFFFFFFF)
        exps[:, e_limbs - 1] |= np.uint32(1 << (hi_bits - 1))
        out.append(SubBatch(N, limbs, size, e_bits, bases, exps))
    return out


class DeviceWorkload:
    """Sub-batches uploaded to the GPU (interleaved) with output buffers.

    The library runs on the context's own (non-blocking) stream, PyTorch's
    fills and copies on PyTorch's stream: every hand-over between the two is
    bracketed by a device synchronisation.
    """

    def __init__(self, ctx: Context, subs: list[SubBatch]):
        import torch

        self.ctx = ctx
        self.subs = subs
        dev = torch.device(f"cuda:{ctx.device}")
        self.mods, self.desc = [], []
        self._tensors = []
        for s in subs:
            stride = stride_for(s.count)
            mod = Modulus(ctx, s.modulus, s.limbs)
            base = self._to_il(torch, dev, s.bases, s.limbs, s.count, stride)
            e_limbs = s.exps.shape[1]
            exps = self._to_il(torch, dev, s.exps, e_limbs, s.count, stride)
            out = torch.zeros((s.limbs, stride), dtype=torch.int32, device=dev)
            self.mods.append(mod)
            self._tensors.append((base, exps, out))
            self.desc.append({"mod": mod, "base": base, "exps": exps, "out": out, "count": s.count,
                              "stride": stride, "e_limbs": e_limbs, "e_bits": s.e_bits})
        torch.cuda.synchronize(dev)

    def _to_il(self, torch, dev, aos: np.ndarray, limbs: int, count: int, stride: int):
        src = torch.from_numpy(np.ascontiguousarray(aos).view(np.int32).copy()).to(dev)
        il = torch.zeros((limbs, stride), dtype=torch.int32, device=dev)
        torch.cuda.synchronize(dev)
        self.ctx.aos_to_il(src, il, count, limbs, stride)
        self.ctx.sync()
        return il

    def result(self, i: int) -> np.ndarray:
        """(count, limbs) uint32 output of sub-batch i."""
        import torch

        d = self.desc[i]
        aos = torch.empty((d["count"], self.subs[i].limbs), dtype=torch.int32, device=d["out"].device)
        torch.cuda.synchronize(d["out"].device)
        self.ctx.il_to_aos(d["out"], aos, d["count"], self.subs[i].limbs, d["stride"])
        self.ctx.sync()
        return aos.cpu().numpy().view(np.uint32)

    def clear_outputs(self) -> None:
        import torch

        self.ctx.sync()
        for d in self.desc:
            d["out"].zero_()
        torch.cuda.synchronize(self.desc[0]["out"].device)

    def total_items(self) -> int:
        return sum(s.count for s in self.subs)

    def algorithmic_macs(self) -> int:
        """Fixed convention: (e + ceil(e/6)) modmuls of 2 L^2 MACs per item."""
        return sum(s.count * (s.e_bits + math.ceil(s.e_bits / 6)) * 2 * s.limbs * s.limbs for s in self.subs)

    def spot_check(self, per_sub: int, rng: random.Random, powm=None) -> list:
        """Compare `per_sub` random items of every sub-batch (always including
        the first and last) with x^e mod N; returns the mismatches as
        (sub-batch, item, count, limbs, got-is-zero)."""
        powm = powm or pow
        bad = []
        self.checked = 0
        for i, s in enumerate(self.subs):
            got = self.result(i)
            picks = {0, s.count - 1} | set(rng.sample(range(s.count), min(per_sub, s.count)))
            self.checked += len(picks)
            for j in sorted(picks):
                b = limbs_to_ints(s.bases[j:j + 1])[0]
                e = limbs_to_ints(s.exps[j:j + 1])[0]
                g = limbs_to_ints(got[j:j + 1])[0]
                if g != powm(b, e, s.modulus):
                    bad.append((i, j, s.count, s.limbs, g == 0))
        return bad
