"""Seeded synthetic databases and queries for the benchmark workloads.

Amino acids are drawn i.i.d. from the Robinson & Robinson (1991) background
frequencies (the composition NCBI BLAST uses); nucleotides uniformly from
ACGT.  Lognormal DB lengths use median 300 (mu = ln 300) and sigma = 0.6,
rounded and clamped to [1, 35000].  Planted homologs are query fragments
mutated to an identity drawn from U(0.60, 0.95) (substitutions drawn from the
background) with indel events at 2 % per residue, geometric lengths of mean 3,
insertions and deletions equally likely.
"""
import numpy as np

AA = "ARNDCQEGHILKMFPSTWYV"
# Robinson & Robinson (1991), Nucleic Acids Res. 19:2599
AA_FREQ = np.array([0.07805, 0.05129, 0.04487, 0.05364, 0.01925, 0.04264, 0.06295, 0.07377, 0.02199, 0.05142,
                    0.09019, 0.05744, 0.02243, 0.03856, 0.05203, 0.07120, 0.05841, 0.01330, 0.03216, 0.06441])
AA_FREQ = AA_FREQ / AA_FREQ.sum()
NT = "ACGT"


def _alphabet(kind):
    return (np.frombuffer(AA.encode(), np.uint8), AA_FREQ) if kind == "aa" else \
        (np.frombuffer(NT.encode(), np.uint8), None)


def random_residues(rng, kind, n):
    alpha, p = _alphabet(kind)
    return alpha[rng.choice(len(alpha), size=n, p=p)]


def lognormal_lengths(rng, n, median=300.0, sigma=0.6, cap=35000):
    return np.clip(np.rint(rng.lognormal(np.log(median), sigma, size=n)), 1, cap).astype(np.int64)


def mutate(rng, kind, seq, identity, indel_rate=0.02, indel_mean=3.0):
    """seq: uint8 array of ASCII residues.  Substitutes each residue with
    probability 1 - identity, then applies indel events at indel_rate per
    residue: an insertion puts a random run before the residue, a deletion
    removes a run starting at it (events inside a deleted run are dropped)."""
    alpha, p = _alphabet(kind)
    s = seq.copy()
    sub = rng.random(len(s)) > identity
    s[sub] = alpha[rng.choice(len(alpha), size=int(sub.sum()), p=p)]
    events = np.nonzero(rng.random(len(s)) < indel_rate)[0]
    if len(events) == 0:
        return s
    pieces, prev = [], 0
    for pos in events.tolist():
        if pos < prev:
            continue
        ln = int(rng.geometric(1.0 / indel_mean))
        pieces.append(s[prev:pos])
        if rng.random() < 0.5:
            pieces.append(random_residues(rng, kind, ln))
            prev = pos
        else:
            prev = pos + ln
    pieces.append(s[prev:])
    out = np.concatenate(pieces)
    return out if len(out) else s[:1]


def make_db(seed, n, kind="aa", lengths="lognormal", queries=(), plant_frac=0.0, fragment=True,
            normal=(200, 50, 100, 300), uniform=(40, 600)):
    """Returns (list of ASCII byte strings, planted indices)."""
    rng = np.random.default_rng(seed)
    if lengths == "lognormal":
        lens = lognormal_lengths(rng, n)
    elif lengths == "normal":
        mu, sd, lo, hi = normal
        lens = np.clip(np.rint(rng.normal(mu, sd, size=n)), lo, hi).astype(np.int64)
    else:
        lens = rng.integers(uniform[0], uniform[1] + 1, size=n)
    total = int(lens.sum())
    res = random_residues(rng, kind, total)
    offs = np.concatenate([[0], np.cumsum(lens)])
    seqs = [res[offs[i]:offs[i + 1]].tobytes() for i in range(n)]
    planted = []
    if queries and plant_frac > 0:
        idx = rng.choice(n, size=int(round(plant_frac * n)), replace=False)
        for k, i in enumerate(sorted(idx.tolist())):
            q = np.frombuffer(queries[k % len(queries)], np.uint8)
            if fragment and len(q) > 8:
                ln = int(rng.integers(max(4, len(q) // 2), len(q) + 1))
                st = int(rng.integers(0, len(q) - ln + 1))
                q = q[st:st + ln]
            seqs[i] = mutate(rng, kind, q, float(rng.uniform(0.60, 0.95))).tobytes()
            planted.append(i)
    return seqs, planted


def make_query(seed, length, kind="aa"):
    rng = np.random.default_rng(seed)
    return random_residues(rng, kind, length).tobytes()


def write_fasta(path, seqs):
    with open(path, "wb") as f:
        parts = []
        for i, s in enumerate(seqs):
            parts.append(b">%d\n" % i)
            parts.append(s)
            parts.append(b"\n")
            if len(parts) > 30000:
                f.write(b"".join(parts))
                parts = []
        f.write(b"".join(parts))


# ---- the benchmark configurations ----
def config_c2(scale=1.0):
    """SW 8-bit cascade: 200,000 lognormal protein sequences, 5 % planted query
    fragments, queries of 50, 500 and 5,000 residues."""
    queries = [make_query(0x5EED0002 + i, L) for i, L in enumerate((50, 500, 5000))]
    db, planted = make_db(0x5EED0002, int(200000 * scale), "aa", "lognormal", queries, 0.05)
    return queries, db, planted


def config_c1():
    """Plain SW, 16-bit lanes: query 120, 2,000 sequences of uniform length 40-600."""
    queries = [make_query(0x5EED0001, 120)]
    db, _ = make_db(0x5EED0001, 2000, "aa", "uniform")
    return queries, db, []


def config_c3(scale=1.0):
    """NW DNA, both strands: query 250 nt, 1,000,000 sequences of normal length
    (200, 50) clipped to 100-300, 1 % planted mutated copies of the query (half of
    them reverse-complemented)."""
    q = make_query(0x5EED0003, 250, "nt")
    db, planted = make_db(0x5EED0003, int(1000000 * scale), "nt", "normal", [q], 0.01, fragment=False)
    comp = bytes.maketrans(b"ACGT", b"TGCA")
    for k, i in enumerate(planted):
        if k % 2:
            db[i] = db[i].translate(comp)[::-1]
    return [q], db, planted


def log_uniform_lengths(rng, n, lo, hi):
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=n))).astype(np.int64)


def config_c4(scale=1.0, nq=256):
    """Batched SW: 256 queries with lengths log-uniform in [30, 2000], 500,000
    lognormal protein sequences (cap 35,000), 2 % planted homologs of the
    queries (round robin)."""
    rng = np.random.default_rng(0x5EED0004)
    lens = log_uniform_lengths(rng, nq, 30, 2000)
    queries = [make_query(0x5EED0400 + i, int(L)) for i, L in enumerate(lens)]
    db, planted = make_db(0x5EED0004, int(500000 * scale), "aa", "lognormal", queries, 0.02)
    return queries, db, planted


def config_c5(scale=1.0):
    """Served SW/NW mix: 1,000,000 lognormal protein sequences (cap 35,000)."""
    db, _ = make_db(0x5EED0005, int(1000000 * scale), "aa", "lognormal")
    return db


def c5_requests(n, rate, seed=0x5EED0500):
    """n requests of the served mix: arrival times of a Poisson stream of the
    given rate (per second), 80 % SW / 20 % NW, query lengths from the mixture
    70 % uniform [50, 400] and 30 % log-uniform [400, 5000]."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
    kinds = np.where(rng.random(n) < 0.8, "sw", "nw")
    short = rng.random(n) < 0.7
    lens = np.where(short, rng.integers(50, 401, size=n), log_uniform_lengths(rng, n, 400, 5000))
    queries = [make_query(seed + 1 + i, int(L)) for i, L in enumerate(lens)]
    return [dict(t=float(arrivals[i]), kind=str(kinds[i]), query=queries[i]) for i in range(n)]
