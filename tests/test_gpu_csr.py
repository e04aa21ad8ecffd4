"""fixed_probability generated on the device, compared with the reference's own
CSRs (tests/golden/csr_small.npz in full, tests/golden/csr.json by FNV-1a hash,
count and first row).  Bit-exact is the bar: the CSR is integer data."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
import spice2_amd as S

pytestmark = pytest.mark.gpu

G = O.ROOT / "tests" / "golden"
CSR = json.loads((G / "csr.json").read_text())


def seed_of(rec):
    return O.seed_chain(rec["seeds"], rec["advance"] + 1)[rec["advance"]]


def generate(src, dst, p, seed):
    with S.Context(max_delay=1) as c:
        a = c.population(src, 0)
        b = c.population(dst, 1)
        pr = c.fixed_probability(a, b, p, seed)
        off, nb = c.download(pr, src)
        st = c.stats()
    return off, nb, st


def test_small_csrs_in_full():
    z = np.load(G / "csr_small.npz")
    for i, rec in enumerate(CSR["full"]):
        if "error" in rec:
            continue
        print("case", rec, flush=True)
        off, nb, st = generate(rec["src"], rec["dst"], rec["p"], seed_of(rec))
        assert (off == z[f"off{i}"]).all(), rec
        assert (nb == z[f"nb{i}"]).all(), rec


def test_small_csrs_exact_row_path(monkeypatch):
    """Every row decided by the exact interval evaluation gives the same CSRs."""
    monkeypatch.setenv("SPICE_CSR_EXACT_ROWS", "1")
    z = np.load(G / "csr_small.npz")
    for i, rec in enumerate(CSR["full"]):
        if "error" in rec or rec["src"] * rec["dst"] > 5e5:
            continue
        print("case", rec, flush=True)
        off, nb, _ = generate(rec["src"], rec["dst"], rec["p"], seed_of(rec))
        assert (off == z[f"off{i}"]).all() and (nb == z[f"nb{i}"]).all(), rec


def test_brunel_family_hashes():
    for rec in CSR["hashed"]:
        if rec["src"] * rec["dst"] > 3e8:
            continue
        print("case", rec["case"], flush=True)
        off, nb, st = generate(rec["src"], rec["dst"], rec["p"], seed_of(rec))
        assert len(nb) == rec["count"], rec["case"]
        assert nb[: off[1]][:8].tolist() == [v for v in rec["row0"] if v >= 0], rec["case"]
        assert f"{O.csr_hash(off, nb):016x}" == rec["fnv"], rec["case"]


def test_empty_source_is_a_precondition():
    with S.Context(max_delay=1) as c:
        a = c.population(0, 0)
        b = c.population(10, 1)
        with pytest.raises(S.SpicePrecondition, match="offsets.size"):
            c.fixed_probability(a, b, 0.1, (1, 2))


def test_probability_out_of_range_is_a_precondition():
    with S.Context(max_delay=1) as c:
        a = c.population(5, 0)
        b = c.population(10, 1)
        with pytest.raises(S.SpicePrecondition, match="0 <= p && p <= 1"):
            c.fixed_probability(a, b, 1.5, (1, 2))


def test_random_cases_against_oracle():
    """Seeds and shapes the golden set does not hold, checked against the oracle
    (which is itself pinned to the reference)."""
    rs = np.random.default_rng(99)
    for _ in range(12):
        src, dst = int(rs.integers(1, 3000)), int(rs.integers(1, 20000))
        p = float(rs.choice([0.0005, 0.003, 0.02, 0.1, 0.35, 0.8, 1.0]))
        seed = (int(rs.integers(0, 2**63)), int(rs.integers(0, 2**63)))
        print("case", src, dst, p, seed, flush=True)
        off, nb, _ = generate(src, dst, p, seed)
        o_off, o_nb = O.fp_generate(src, dst, p, seed)
        assert (off == o_off).all() and (nb == o_nb).all()


@pytest.mark.slow
def test_n500k_projection_hashes():
    """The six N=500k projections (2.5e9 synapses), when their hashes are recorded."""
    recs = [r for r in CSR["hashed"] if r["case"].startswith("c1.")]
    if not recs:
        pytest.skip("N=500k golden hashes not recorded")
    for rec in recs:
        print("case", rec["case"], flush=True)
        off, nb, st = generate(rec["src"], rec["dst"], rec["p"], seed_of(rec))
        assert len(nb) == rec["count"], rec["case"]
        assert f"{O.csr_hash(off, nb):016x}" == rec["fnv"], rec["case"]
        del off, nb
