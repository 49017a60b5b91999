"""The benchmark workloads: seeded synthetic genomes and reads of the five
mapping configurations, shared by bench.py and the GPU parity tests so that
both run exactly the same inputs (the reference ships no data)."""
import swmapper_amd as swm

GENOME_SEED = 20240601
READS_SEED = 777

CONFIGS = {
    "small": dict(n=4_000_000, kind=swm.GENOME_UNIFORM, frac=0.0, reads=100_000, lmin=150, lmax=150,
                  err=(0.025, 0.025), e=5,
                  desc="4 Mbp uniform, 12-mer index; 100k x 150 bp reads, 2% sub + 0.5% indel"),
    "repeat5": dict(n=256_000_000, kind=swm.GENOME_REPEAT_FAMILIES, frac=0.05, reads=1_000_000, lmin=150,
                    lmax=150, err=(0.025, 0.025), e=5,
                    desc="256 Mbp uniform + 5% repeat families (300 bp x 500 copies, 10% div), 12-mer index; "
                         "1M x 150 bp reads, 2% sub + 0.5% 1-3 bp indel"),
    "uneven": dict(n=256_000_000, kind=swm.GENOME_REPEAT_FAMILIES, frac=0.05, reads=1_000_000, lmin=50, lmax=250,
                   err=(0.01, 0.04), e=5,
                   desc="256 Mbp + 5% repeat families, 12-mer index; 1M reads, length U[50,250], error U[1%,4%] "
                        "(80% sub / 20% 1-3 bp indel)"),
    "skew": dict(n=256_000_000, kind=swm.GENOME_TANDEM, frac=0.30, reads=500_000, lmin=150, lmax=150,
                 err=(0.025, 0.025), e=5,
                 desc="256 Mbp + 30% tandem repeats (period U[2,50]), 12-mer index; 500k x 150 bp reads, "
                      "2% sub + 0.5% indel"),
}


def genome(name):
    c = CONFIGS[name]
    return swm.synth_genome(c["n"], seed=GENOME_SEED, kind=c["kind"], frac=c["frac"])


def reads(name, g, n=None):
    """(reads[n, 256], lens[n]) of the configuration (its first n reads)."""
    c = CONFIGS[name]
    r, l, _, _ = swm.synth_reads(g, c["reads"] if n is None else n, len_min=c["lmin"], len_max=c["lmax"],
                                 err_min=c["err"][0], err_max=c["err"][1], indel_share=0.2, seed=READS_SEED,
                                 stride=256)
    return r, l
