"""Seeded inputs shared by the tests (genomes and reads from the library's
workload generator, which runs on the host CPU)."""
import numpy as np

import swmapper_amd as swm


def genome(n, seed=1, kind=swm.GENOME_UNIFORM, frac=0.0):
    return swm.synth_genome(n, seed=seed, kind=kind, frac=frac)


def reads(g, n, len_min=150, len_max=150, err=(0.025, 0.025), seed=7, stride=256):
    return swm.synth_reads(g, n, len_min=len_min, len_max=len_max, err_min=err[0], err_max=err[1],
                           indel_share=0.2, seed=seed, stride=stride)


def exact_reads(g, starts, L, strands, stride=256):
    """Error-free reads copied from the genome (reverse complement on strand 1)."""
    comp = np.zeros(256, dtype=np.uint8)
    for a, b in zip(b"ACGT", b"TGCA"):
        comp[a] = b
    out = np.zeros((len(starts), stride), dtype=np.uint8)
    for i, (s, st) in enumerate(zip(starts, strands)):
        frag = g[s:s + L]
        out[i, :L] = comp[frag[::-1]] if st else frag
    return out, np.full(len(starts), L, dtype=np.uint32)
