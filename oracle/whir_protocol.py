"""CPU restatement of the WHIR protocol layer: parameter selection, prover and
verifier, over the oracle kernels of oracle/whir_oracle.c.

TEST INFRASTRUCTURE ONLY (imported by tests/, smoke() and bench.py's
cpu_baseline leg as the checker; never by the product).

Restated from the reference (WizardOfMenlo/whir):
  Config.new                  src/protocols/whir/config.rs:16-152
  IrsConfig.new               src/protocols/irs_commit.rs:125-196, 579-630
  soundness terms             src/protocols/irs_commit.rs:240-288
  commit                      src/protocols/irs_commit.rs:292-350 (+ matrix_commit.rs:211-232,
                              merkle_tree.rs:82-127)
  open                        src/protocols/irs_commit.rs:385-432, merkle_tree.rs:145-186
  challenge_indices           src/protocols/challenge_indices.rs:6-50
  geometric_challenge         src/protocols/geometric_challenge.rs:10-24
  sumcheck prove / verify     src/protocols/sumcheck.rs:69-216 (mask_length = 0)
  proof of work               src/protocols/proof_of_work.rs:68-192
  prove                       src/protocols/whir/prover.rs:50-316 (MLE points; several
                              witnesses of batch_size vectors each)
  verify                      src/protocols/whir/verifier.rs:45-264

Transcript: the reference drives spongefish (git fcc277f8, not vendored here), whose
duplex construction and codecs cannot be pinned in this environment.  Transcript
below is this project's own Keccak-f[1600] duplex with the typed operations the
reference uses (prover field / hash / u64 messages, verifier field / byte
challenges, hash and field-vector hints).  The GPU prover implements the same
spec, so proof bytes are compared byte for byte; parity of these bytes with
spongefish itself is unpinned.
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from oracle import oracle as O

P = O.P
LOG2_10 = math.log2(10.0)
U64_MAX = 2**64 - 1


# ---------------------------------------------------------------- transcript
class Transcript:
    """Keccak-f[1600] duplex, rate 136 bytes.

    absorb: XOR bytes into the rate, permuting when it is full; squeeze: permute on
    the switch from absorbing, then read the rate, permuting when it is used up.
    Field challenges: 24 squeezed bytes per base coefficient, little-endian, mod p.
    """

    RATE = 136

    def __init__(self, protocol_id: bytes, session_id: bytes, instance: bytes = b""):
        self.state = np.zeros(25, dtype=np.uint64)
        self.pos = 0
        self.squeezing = False
        self.narg = bytearray()
        self.hints = bytearray()
        self._absorb(bytes(protocol_id) + bytes(session_id) + bytes(instance))

    def _permute(self):
        O.lib().oracle_keccak_f(self.state.ctypes.data_as(O.u64p))

    def _bytes(self):
        return self.state.view(np.uint8)

    def _absorb(self, data: bytes):
        self.squeezing = False
        for b in data:
            if self.pos == self.RATE:
                self._permute()
                self.pos = 0
            self._bytes()[self.pos] ^= b
            self.pos += 1

    def _squeeze(self, n: int) -> bytes:
        if not self.squeezing:
            self._permute()
            self.pos = 0
            self.squeezing = True
        out = bytearray()
        for _ in range(n):
            if self.pos == self.RATE:
                self._permute()
                self.pos = 0
            out.append(int(self._bytes()[self.pos]))
            self.pos += 1
        return bytes(out)

    # typed operations (ProverState, src/transcript/mod.rs:183-267)
    def prover_field(self, D: int, elems):
        e = np.asarray(elems, dtype=np.uint64).reshape(-1)
        assert e.size % D == 0
        b = e.astype("<u8").tobytes()
        self.narg += b
        self._absorb(b)

    def verifier_field(self, D: int) -> np.ndarray:
        return np.array([int.from_bytes(self._squeeze(24), "little") % P for _ in range(D)], dtype=np.uint64)

    def verifier_bytes(self, n: int) -> bytes:
        return self._squeeze(n)

    def prover_hash(self, h: bytes):
        assert len(h) == 32
        self.narg += h
        self._absorb(h)

    def prover_u64(self, v: int):
        b = struct.pack("<Q", v)
        self.narg += b
        self._absorb(b)

    def hint_hash(self, h: bytes):
        self.hints += h

    def hint_field_vec(self, D: int, elems):
        e = np.asarray(elems, dtype=np.uint64).reshape(-1)
        self.hints += struct.pack("<Q", e.size // D) + e.astype("<u8").tobytes()


class TranscriptReader:
    """Verifier side of Transcript: replays the same duplex over a proof."""

    def __init__(self, protocol_id: bytes, session_id: bytes, narg: bytes, hints: bytes, instance: bytes = b""):
        self.t = Transcript(protocol_id, session_id, instance)
        self.narg = bytes(narg)
        self.hints = bytes(hints)
        self.np = 0
        self.hp = 0

    def _take(self, n):
        if self.np + n > len(self.narg):
            raise VerificationError("narg exhausted")
        b = self.narg[self.np:self.np + n]
        self.np += n
        return b

    def prover_field(self, D: int, count: int = 1) -> np.ndarray:
        b = self._take(8 * D * count)
        self.t._absorb(b)
        v = np.frombuffer(b, dtype="<u8").astype(np.uint64)
        if (v >= np.uint64(P)).any():
            raise VerificationError("non-canonical field element")
        return v

    def verifier_field(self, D: int) -> np.ndarray:
        return self.t.verifier_field(D)

    def verifier_bytes(self, n: int) -> bytes:
        return self.t.verifier_bytes(n)

    def prover_hash(self) -> bytes:
        b = self._take(32)
        self.t._absorb(b)
        return b

    def prover_u64(self) -> int:
        b = self._take(8)
        self.t._absorb(b)
        return struct.unpack("<Q", b)[0]

    def hint_hash(self) -> bytes:
        if self.hp + 32 > len(self.hints):
            raise VerificationError("hints exhausted")
        b = self.hints[self.hp:self.hp + 32]
        self.hp += 32
        return b

    def hint_field_vec(self, D: int) -> np.ndarray:
        if self.hp + 8 > len(self.hints):
            raise VerificationError("hints exhausted")
        n = struct.unpack("<Q", self.hints[self.hp:self.hp + 8])[0]
        self.hp += 8
        nb = 8 * D * n
        if self.hp + nb > len(self.hints):
            raise VerificationError("hints exhausted")
        v = np.frombuffer(self.hints[self.hp:self.hp + nb], dtype="<u8").astype(np.uint64)
        self.hp += nb
        return v

    def check_eof(self):
        if self.np != len(self.narg) or self.hp != len(self.hints):
            raise VerificationError("trailing proof bytes")


class VerificationError(Exception):
    pass


def vassert(cond, msg="verification failed"):
    if not cond:
        raise VerificationError(msg)


# ---------------------------------------------------------------- parameters
@dataclass
class ProtocolParameters:
    """src/parameters.rs:8-24"""
    security_level: int = 100
    pow_bits: int = 20
    initial_folding_factor: int = 4
    folding_factor: int = 4
    unique_decoding: bool = False
    starting_log_inv_rate: int = 1
    batch_size: int = 1
    hash_engine: int = O.ENGINE_BLAKE3


def field_size_bits(D: int) -> float:
    """FieldWithSize::field_size_bits, src/algebra/fields.rs:14-30 (f64 of the modulus)."""
    return math.log2(float(P)) * D


def pow_threshold(bits: float) -> int:
    return O.pow_threshold(bits)


def pow_difficulty(threshold: int) -> float:
    """proof_of_work::difficulty, src/protocols/proof_of_work.rs:41-43"""
    return 64.0 - math.log2(float(threshold))


@dataclass
class IrsConfig:
    num_vectors: int
    vector_size: int
    codeword_length: int
    interleaving_depth: int
    johnson_slack: float
    in_domain_samples: int
    out_domain_samples: int
    target_degree: int
    mask_length: int = 0
    deduplicate_in_domain: bool = False

    @staticmethod
    def new(security_target: float, unique_decoding: bool, num_vectors: int, vector_size: int,
            interleaving_depth: int, rate: float, target_degree: int) -> "IrsConfig":
        """irs_commit::Config::new, src/protocols/irs_commit.rs:125-196"""
        assert vector_size % interleaving_depth == 0
        assert 0.0 < rate <= 1.0
        message_length = vector_size // interleaving_depth
        codeword_length = int(math.ceil(message_length / rate))
        rate = message_length / codeword_length
        johnson_slack = 0.0 if unique_decoding else math.sqrt(rate) / 20.0
        if unique_decoding:
            ood = 0
        else:
            list_size = 1.0 / (2.0 * johnson_slack * math.sqrt(rate))
            ood = num_ood_samples(False, security_target, field_size_bits(target_degree), list_size, vector_size)
        per_sample = (1.0 + rate) / 2.0 if unique_decoding else math.sqrt(rate) + johnson_slack
        ind = int(math.ceil(security_target / (-math.log2(per_sample))))
        return IrsConfig(num_vectors, vector_size, codeword_length, interleaving_depth, johnson_slack, ind, ood,
                         target_degree)

    def num_messages(self):
        return self.interleaving_depth * self.num_vectors

    def message_length(self):
        return self.vector_size // self.interleaving_depth

    def masked_message_length(self):
        return self.message_length() + self.mask_length

    def rate(self):
        return self.masked_message_length() / self.codeword_length

    def unique_decoding(self):
        return self.out_domain_samples == 0 and self.johnson_slack == 0.0

    def list_size(self):
        if self.unique_decoding():
            return 1.0
        return 1.0 / (2.0 * self.johnson_slack * math.sqrt(self.rate()))

    def rbr_queries(self):
        per_sample = (1.0 + self.rate()) / 2.0 if self.unique_decoding() else math.sqrt(self.rate()) + self.johnson_slack
        return self.in_domain_samples * -math.log2(per_sample)

    def rbr_soundness_fold_prox_gaps(self):
        log_field = field_size_bits(self.target_degree)
        log_inv_rate = -math.log2(self.rate())
        log_k = math.log2(float(self.masked_message_length()))
        if self.unique_decoding():
            error = log_k + log_inv_rate
        else:
            error = 7.0 * LOG2_10 + 3.5 * log_inv_rate + 2.0 * log_k
        return log_field - error


def num_ood_samples(unique_decoding, security_target, fsb, list_size, degree):
    """src/protocols/irs_commit.rs:584-601"""
    if unique_decoding:
        return 0
    l_choose_2 = list_size * (list_size - 1.0) / 2.0
    log_per_sample = fsb - math.log2(float(degree - 1))
    assert log_per_sample > 0.0
    return int(max(math.ceil((security_target + math.log2(l_choose_2)) / log_per_sample), 1.0))


@dataclass
class SumcheckConfig:
    initial_size: int
    round_pow: int  # threshold
    num_rounds: int

    def final_size(self):
        if self.initial_size == 0 or self.num_rounds == 0:
            return self.initial_size
        return next_pow2(self.initial_size) >> self.num_rounds


@dataclass
class RoundConfig:
    irs: IrsConfig
    sumcheck: SumcheckConfig
    pow: int  # threshold


@dataclass
class WhirConfig:
    D: int
    hash_engine: int
    initial: IrsConfig
    initial_sumcheck: SumcheckConfig
    initial_skip_pow: int
    rounds: List[RoundConfig]
    final_sumcheck: SumcheckConfig
    final_pow: int

    @staticmethod
    def new(size: int, p: ProtocolParameters, D: int) -> "WhirConfig":
        """whir::Config::new, src/protocols/whir/config.rs:16-152"""
        assert size > 0 and size & (size - 1) == 0
        pw = lambda bits: pow_threshold(bits)  # noqa: E731
        security_level = float(p.security_level)
        protocol_security_level = float(max(p.security_level - p.pow_bits, 0))
        fsb = field_size_bits(D)
        log_inv_rate = p.starting_log_inv_rate
        num_variables = size.bit_length() - 1
        initial = IrsConfig.new(protocol_security_level, p.unique_decoding, p.batch_size, size,
                                1 << p.initial_folding_factor, 0.5 ** p.starting_log_inv_rate, D)
        prox = initial.rbr_soundness_fold_prox_gaps()
        sumcheck_error = fsb - math.log2(initial.list_size()) - 1.0
        starting_folding_pow_bits = max(security_level - min(prox, sumcheck_error), 0.0)
        initial_skip_pow_bits = max(security_level - (prox + math.log2(float(p.initial_folding_factor))), 0.0)
        rounds = []
        rnd = 0
        in_domain_samples = initial.in_domain_samples
        query_error = initial.rbr_queries()
        num_variables -= p.initial_folding_factor
        while num_variables >= p.folding_factor:
            rff = p.initial_folding_factor if rnd == 0 else p.folding_factor
            next_rate = log_inv_rate + (rff - 1)
            irs = IrsConfig.new(protocol_security_level, p.unique_decoding, 1, 1 << num_variables,
                                1 << p.folding_factor, 0.5 ** next_rate, D)
            log_list = math.log2(irs.list_size())
            count = irs.out_domain_samples + in_domain_samples
            combination_error = fsb - (math.log2(float(count)) + log_list + 1.0)
            pow_bits = max(0.0, security_level - min(query_error, combination_error))
            fprox = irs.rbr_soundness_fold_prox_gaps()
            fsum = fsb - (log_list + 1.0)
            folding_pow_bits = max(security_level - min(fprox, fsum), 0.0)
            rounds.append(RoundConfig(irs, SumcheckConfig(1 << num_variables, pw(folding_pow_bits), p.folding_factor),
                                      pw(pow_bits)))
            rnd += 1
            num_variables -= p.folding_factor
            log_inv_rate = next_rate
            in_domain_samples = irs.in_domain_samples
            query_error = irs.rbr_queries()
        rbr_error = rounds[-1].irs.rbr_queries() if rounds else initial.rbr_queries()
        final_pow_bits = max(0.0, security_level - rbr_error)
        final_folding_pow_bits = max(0.0, security_level - fsb + 1.0)
        return WhirConfig(D, p.hash_engine, initial,
                          SumcheckConfig(size, pw(starting_folding_pow_bits), p.initial_folding_factor),
                          pw(initial_skip_pow_bits), rounds,
                          SumcheckConfig(1 << num_variables, pw(final_folding_pow_bits), num_variables),
                          pw(final_pow_bits))

    def disable_pow(self):
        """whir::Config::disable_pow, src/protocols/whir/mod.rs:107-118"""
        self.initial_sumcheck.round_pow = U64_MAX
        self.initial_skip_pow = U64_MAX
        for r in self.rounds:
            r.sumcheck.round_pow = U64_MAX
            r.pow = U64_MAX
        self.final_sumcheck.round_pow = U64_MAX
        self.final_pow = U64_MAX

    def describe(self) -> dict:
        def irs(c: IrsConfig):
            return {"vector_size": c.vector_size, "codeword_length": c.codeword_length,
                    "interleaving_depth": c.interleaving_depth, "num_vectors": c.num_vectors,
                    "in_domain_samples": c.in_domain_samples, "out_domain_samples": c.out_domain_samples}
        return {
            "degree": self.D, "hash": self.hash_engine,
            "initial": irs(self.initial),
            "initial_sumcheck": [self.initial_sumcheck.initial_size, self.initial_sumcheck.num_rounds,
                                 self.initial_sumcheck.round_pow],
            "initial_skip_pow": self.initial_skip_pow,
            "rounds": [{"irs": irs(r.irs), "sumcheck": [r.sumcheck.initial_size, r.sumcheck.num_rounds,
                                                         r.sumcheck.round_pow], "pow": r.pow} for r in self.rounds],
            "final_sumcheck": [self.final_sumcheck.initial_size, self.final_sumcheck.num_rounds,
                               self.final_sumcheck.round_pow],
            "final_pow": self.final_pow,
        }


def next_pow2(n):
    p = 1
    while p < n:
        p <<= 1
    return p


# ---------------------------------------------------------------- helpers over the oracle kernels
def ext(D, x):
    return np.asarray(x, dtype=np.uint64).reshape(-1)


def e_one(D):
    v = np.zeros(D, dtype=np.uint64)
    v[0] = 1
    return v


def e_mul(D, a, b):
    return O.mul(D, ext(D, a), ext(D, b)).reshape(-1)


def e_add(D, a, b):
    return O.add(D, ext(D, a), ext(D, b)).reshape(-1)


def e_sub(D, a, b):
    return O.sub(D, ext(D, a), ext(D, b)).reshape(-1)


def geometric_sequence(D, x, n):
    out = []
    cur = e_one(D)
    for _ in range(n):
        out.append(cur)
        cur = e_mul(D, cur, x)
    return out


def geometric_challenge(t, D, count):
    """src/protocols/geometric_challenge.rs:10-24"""
    if count == 0:
        return []
    if count == 1:
        return [e_one(D)]
    x = t.verifier_field(D)
    return geometric_sequence(D, x, count)


def challenge_indices(t, num_leaves, count, deduplicate=False):
    """src/protocols/challenge_indices.rs:6-50 (one u8 challenge per byte)."""
    if count == 0:
        return []
    assert num_leaves & (num_leaves - 1) == 0
    if num_leaves == 1:
        return [0] if deduplicate else [0] * count
    size_bytes = ((num_leaves.bit_length() - 1) + 7) // 8
    entropy = b"".join(t.verifier_bytes(1) for _ in range(count * size_bytes))
    idx = O.bytes_to_indices(entropy, count, num_leaves)
    if deduplicate:
        idx = sorted(set(idx))
    return idx


def eq_weights(D, point_list):
    k = len(point_list)
    pt = np.concatenate(point_list) if k else np.zeros(0, dtype=np.uint64)
    return O.eval_eq(D, np.zeros((1 << k) * D, dtype=np.uint64), pt, e_one(D)).reshape(-1, D)


def pow_prove(t, engine, threshold):
    """proof_of_work::Config::prove, src/protocols/proof_of_work.rs:68-166"""
    if threshold == U64_MAX:
        return
    ch = t.verifier_bytes(32)
    nonce = O.pow_search(engine, ch, threshold, 0, 1 << 40)
    assert nonce is not None, "Proof of Work failed to solve."
    t.prover_u64(nonce)


def pow_verify(r, engine, threshold):
    if threshold == U64_MAX:
        return
    ch = r.verifier_bytes(32)
    nonce = r.prover_u64()
    h = O.hash_many(engine, 64, ch + struct.pack("<Q", nonce) + b"\0" * 24, 1)
    vassert(int.from_bytes(h[:8], "little") <= threshold, "pow")


# ---------------------------------------------------------------- IRS commit / open
@dataclass
class Witness:
    source_degree: int
    matrix: np.ndarray   # row-major codeword_length x num_messages (elements of source_degree)
    nodes: bytes
    ood_points: list
    ood_values: list     # point-major: out_domain_samples x num_vectors
    num_vectors: int = 1


def irs_commit(t, cfg: IrsConfig, engine: int, vectors: list, Ds: int, D: int) -> Witness:
    """irs_commit::Config::commit, src/protocols/irs_commit.rs:292-350"""
    L = cfg.message_length()
    msgs = []
    for v in vectors:
        v = ext(Ds, v)
        for j in range(cfg.interleaving_depth):
            msgs.append(v[j * L * Ds:(j + 1) * L * Ds])
    m = np.stack(msgs) if msgs else np.zeros((0, 0), dtype=np.uint64)
    matrix = O.interleaved_encode(Ds, m, cfg.codeword_length)
    leaves = O.hash_rows(engine, Ds, matrix.reshape(-1), cfg.codeword_length, cfg.num_messages())
    nodes = O.merkle_commit(engine, leaves)
    t.prover_hash(nodes[-32:])
    pts = [t.verifier_field(D) for _ in range(cfg.out_domain_samples)]
    vals = []
    for x in pts:
        for v in vectors:
            val = O.univariate_evaluate(Ds, D, ext(Ds, v), x)
            t.prover_field(D, val)
            vals.append(val)
    return Witness(Ds, matrix, nodes, pts, vals, len(vectors))


def irs_open(t, cfg: IrsConfig, w: Witness):
    """irs_commit::Config::open, src/protocols/irs_commit.rs:385-432, for one witness;
    returns (indices, points, rows)."""
    return irs_open_many(t, cfg, [w])


def irs_open_many(t, cfg: IrsConfig, witnesses: list):
    """irs_commit::Config::open over several witnesses (irs_commit.rs:385-432): one
    set of in-domain indices; per witness its opened rows and Merkle path as hints.
    rows[q] is the horizontal concatenation of every witness's row idx[q]."""
    idx = challenge_indices(t, cfg.codeword_length, cfg.in_domain_samples, cfg.deduplicate_in_domain)
    pts = [int(x) for x in O.evaluation_points(cfg.masked_message_length(), cfg.codeword_length, idx)] if idx else []
    per = []
    for w in witnesses:
        rows = [w.matrix[i].reshape(-1) for i in idx]
        sub = np.concatenate(rows) if rows else np.zeros(0, dtype=np.uint64)
        t.hint_field_vec(w.source_degree, sub)
        hints = O.merkle_open(w.nodes, cfg.codeword_length, idx)
        for i in range(0, len(hints), 32):
            t.hint_hash(hints[i:i + 32])
        per.append(rows)
    rows = [np.concatenate([r[q] for r in per]) for q in range(len(idx))]
    return idx, pts, rows


def irs_verify(r, cfg: IrsConfig, engine: int, root: bytes, Ds: int):
    """irs_commit::Config::verify, src/protocols/irs_commit.rs:438-484 (one commitment)."""
    return irs_verify_many(r, cfg, engine, [root], Ds)


def irs_verify_many(r, cfg: IrsConfig, engine: int, roots: list, Ds: int):
    """irs_commit::Config::verify over several commitments (irs_commit.rs:438-484);
    rows are horizontally concatenated across the commitments."""
    idx = challenge_indices(r, cfg.codeword_length, cfg.in_domain_samples, cfg.deduplicate_in_domain)
    pts = [int(x) for x in O.evaluation_points(cfg.masked_message_length(), cfg.codeword_length, idx)] if idx else []
    cols = cfg.num_messages()
    per = []
    for root in roots:
        sub = r.hint_field_vec(Ds)
        vassert(sub.size == len(idx) * cols * Ds, "opened rows")
        leaves = O.hash_rows(engine, Ds, sub, len(idx), cols) if idx else b""
        merkle_verify(r, engine, cfg.codeword_length, root, idx,
                      [leaves[32 * i:32 * i + 32] for i in range(len(idx))])
        per.append([sub[k * cols * Ds:(k + 1) * cols * Ds] for k in range(len(idx))])
    rows = [np.concatenate([p[k] for p in per]) for k in range(len(idx))]
    return pts, rows


def merkle_verify(r, engine, num_leaves, root, indices, leaf_hashes):
    """merkle_tree::Config::verify, src/protocols/merkle_tree.rs:188-273"""
    if not indices:
        return
    layers = O.layers_for_size(num_leaves)
    pairs = sorted(zip(indices, leaf_hashes), key=lambda x: x[0])
    for i in range(1, len(pairs)):
        if pairs[i - 1][0] == pairs[i][0]:
            vassert(pairs[i - 1][1] == pairs[i][1], "duplicate leaf mismatch")
    dedup = []
    for p in pairs:
        if not dedup or dedup[-1][0] != p[0]:
            dedup.append(p)
    idx = [p[0] for p in dedup]
    hs = [p[1] for p in dedup]
    for _ in range(layers):
        ni, inp = [], b""
        k = 0
        while k < len(idx):
            a = idx[k]
            if k + 1 < len(idx) and idx[k + 1] == a ^ 1:
                inp += hs[k] + hs[k + 1]
                k += 2
            else:
                h = r.hint_hash()
                inp += (hs[k] + h) if a % 2 == 0 else (h + hs[k])
                k += 1
            ni.append(a >> 1)
        out = O.hash_many(engine, 64, inp, len(inp) // 64)
        idx, hs = ni, [out[32 * i:32 * i + 32] for i in range(len(ni))]
    vassert(idx == [0] and hs[0] == root, "merkle root")


# ---------------------------------------------------------------- sumcheck
def sumcheck_prove(t, cfg: SumcheckConfig, D, engine, a, b, the_sum):
    """sumcheck::Config::prove with no masks, src/protocols/sumcheck.rs:69-162"""
    assert a.size // D == cfg.initial_size
    chals = []
    prev = None
    for _ in range(cfg.num_rounds):
        if prev is not None:
            a = O.fold(D, a, prev)
            b = O.fold(D, b, prev)
        c0, c2 = O.sumcheck_poly(D, a, b)
        c1 = e_sub(D, e_sub(D, e_sub(D, the_sum, c0), c0), c2)
        t.prover_field(D, c0)
        t.prover_field(D, c2)
        pow_prove(t, engine, cfg.round_pow)
        r = t.verifier_field(D)
        chals.append(r)
        the_sum = e_add(D, e_mul(D, e_add(D, e_mul(D, c2, r), c1), r), c0)
        prev = r
    if prev is not None:
        a = O.fold(D, a, prev)
        b = O.fold(D, b, prev)
    return a, b, the_sum, chals


def sumcheck_verify(r, cfg: SumcheckConfig, D, engine, the_sum):
    """sumcheck::Config::verify, src/protocols/sumcheck.rs:165-216"""
    chals = []
    for _ in range(cfg.num_rounds):
        c0 = r.prover_field(D)
        c2 = r.prover_field(D)
        c1 = e_sub(D, e_sub(D, e_sub(D, the_sum, c0), c0), c2)
        pow_verify(r, engine, cfg.round_pow)
        x = r.verifier_field(D)
        chals.append(x)
        the_sum = e_add(D, e_mul(D, e_add(D, e_mul(D, c2, x), c1), x), c0)
    return the_sum, chals


# ---------------------------------------------------------------- WHIR prove / verify
def commit(t, cfg: WhirConfig, vector) -> Witness:
    """whir::Config::commit, src/protocols/whir/mod.rs:79-92 (base-field vector)."""
    return commit_batch(t, cfg, [vector])


def commit_batch(t, cfg: WhirConfig, vectors: list) -> Witness:
    """whir::Config::commit of batch_size base-field vectors in one commitment
    (src/protocols/whir/mod.rs:79-92 -> irs_commit.rs:292-350)."""
    assert len(vectors) == cfg.initial.num_vectors, "commit: batch_size vectors per commitment"
    return irs_commit(t, cfg.initial, cfg.hash_engine, [ext(1, v) for v in vectors], 1, cfg.D)


def mle_evaluate_base(D, vector, point_list):
    """MultilinearExtension evaluate of a base-field vector (mixed_multilinear_extend)."""
    lv = O.lift(D, ext(1, vector))
    return O.multilinear_extend(D, lv, np.concatenate(point_list))


@dataclass
class Covector:
    """Linear form given as an explicit covector (src/algebra/linear_form/covector.rs):
    vector_size extension elements (D words each), evaluated as a dot product."""
    vector: np.ndarray


@dataclass
class Univariate:
    """UnivariateEvaluation linear form (src/algebra/linear_form/univariate_evaluation.rs):
    v -> sum_i v_i x^i at an extension point x (D words)."""
    point: np.ndarray


def form_evaluate(D, form, vector):
    """Evaluate::evaluate of any of the three linear forms on a base-field vector."""
    if isinstance(form, Covector):
        return covector_evaluate(D, form, vector)
    if isinstance(form, Univariate):
        return O.univariate_evaluate(1, D, ext(1, vector), np.asarray(form.point, dtype=np.uint64))
    return mle_evaluate_base(D, vector, form)


def covector_evaluate(D, form: Covector, vector):
    """Evaluate::evaluate of a Covector on a base-field vector (covector.rs:40-45, mixed_dot)."""
    return O.dot(D, np.asarray(form.vector, dtype=np.uint64).reshape(-1), O.lift(D, ext(1, vector)))


def _row_dot(D, weights, row, Ds):
    lrow = O.lift(D, row) if Ds == 1 else row
    return O.dot(D, np.concatenate(weights), lrow)


def _tensor(D, a, b):
    """tensor_product, src/algebra/mod.rs:40-48."""
    return [e_mul(D, x, y) for x in a for y in b]


def prove(t, cfg: WhirConfig, vector, witness: Witness, points: list, evaluations: list):
    """whir::Config::prove, src/protocols/whir/prover.rs:50-316, with one committed
    vector and MultilinearExtension linear forms at `points` (each a list of
    num_variables extension elements)."""
    return prove_batch(t, cfg, [vector], [witness], points, [[e] for e in evaluations])


def prove_batch(t, cfg: WhirConfig, vectors: list, witnesses: list, points: list, evaluations: list,
                check_invariants: bool = False):
    """whir::Config::prove, src/protocols/whir/prover.rs:50-316: vectors committed in
    `witnesses` (batch_size each, in order); linear forms in `points`, each a
    MultilinearExtension point (a list of num_variables elements) or a Covector;
    evaluations[p][j] = form p applied to vector j.  check_invariants
    asserts dot(vector, covector) == the_sum where the reference debug-asserts it."""
    D, eng = cfg.D, cfg.hash_engine
    n = cfg.initial.vector_size
    V = len(vectors)
    assert V == len(witnesses) * cfg.initial.num_vectors and V > 0
    assert len(evaluations) == len(points) and all(len(row) == V for row in evaluations)
    # Complete the OOD matrix with the cross terms (prover.rs:98-125): each
    # witness's OOD rows hold its own vectors; the others are computed and sent.
    oods_points, oods_matrix = [], []
    off = 0
    for w in witnesses:
        B = w.num_vectors
        for i, x in enumerate(w.ood_points):
            row = []
            for j in range(V):
                if off <= j < off + B:
                    row.append(w.ood_values[i * B + j - off])
                else:
                    val = O.univariate_evaluate(1, D, ext(1, vectors[j]), x)
                    t.prover_field(D, val)
                    row.append(val)
            oods_points.append(x)
            oods_matrix.append(row)
        off += B
    vector_rlc = geometric_challenge(t, D, V)
    vec = O.lift(D, ext(1, vectors[0]))
    for c, v in zip(vector_rlc[1:], vectors[1:]):
        vec = O.mixed_scalar_mul_add(D, vec, c, ext(1, v))
    crlc = geometric_challenge(t, D, len(points) + len(oods_points))
    has_constraints = len(crlc) > 0
    form_rlc, oods_rlc = crlc[:len(points)], crlc[len(points):]
    cov = np.zeros(n * D, dtype=np.uint64)
    for rc, pt in zip(form_rlc, points):
        if isinstance(pt, Covector):  # Covector::accumulate, covector.rs:21-23
            v = np.asarray(pt.vector, dtype=np.uint64).reshape(-1)
            cov = O.add(D, cov, O.mul(D, v, np.tile(np.asarray(rc, dtype=np.uint64), n))).reshape(-1)
        elif isinstance(pt, Univariate):  # UnivariateEvaluation::accumulate (geometric_accumulate)
            cov = O.geometric_accumulate(D, cov, np.asarray(rc, dtype=np.uint64),
                                         np.asarray(pt.point, dtype=np.uint64))
        else:
            cov = O.eval_eq(D, cov, np.concatenate(pt), rc)
    the_sum = np.zeros(D, dtype=np.uint64)
    for rc, row in zip(form_rlc, evaluations):
        the_sum = e_add(D, the_sum, e_mul(D, rc, _row_dot(D, vector_rlc, np.concatenate(row), D)))
    if oods_points:
        cov = O.geometric_accumulate(D, cov, np.concatenate(oods_rlc), np.concatenate(oods_points))
    for rc, row in zip(oods_rlc, oods_matrix):
        the_sum = e_add(D, the_sum, e_mul(D, rc, _row_dot(D, vector_rlc, np.concatenate(row), D)))
    if check_invariants and has_constraints:
        assert np.array_equal(O.dot(D, vec, cov), the_sum), "dot(vector, covector) != the_sum"
    if has_constraints:
        vec, cov, the_sum, fr = sumcheck_prove(t, cfg.initial_sumcheck, D, eng, vec, cov, the_sum)
    else:
        fr = [t.verifier_field(D) for _ in range(cfg.initial_sumcheck.num_rounds)]
        pow_prove(t, eng, cfg.initial_skip_pow)
        for f in fr:
            vec = O.fold(D, vec, f)
        cov = np.zeros(cfg.initial_sumcheck.final_size() * D, dtype=np.uint64)
    evaluation_point = list(fr)
    prev_cfg, prev_ws = cfg.initial, list(witnesses)
    for rc in cfg.rounds:
        new_w = irs_commit(t, rc.irs, eng, [vec], D, D)
        pow_prove(t, eng, rc.pow)
        idx, pts, rows = irs_open_many(t, prev_cfg, prev_ws)
        weights = _tensor(D, vector_rlc, list(eq_weights(D, fr)))
        stir_points = list(new_w.ood_points) + [O.lift(D, [p]) for p in pts]
        stir_vals = list(new_w.ood_values)
        for row in rows:
            stir_vals.append(_row_dot(D, weights, row, prev_ws[0].source_degree))
        srlc = geometric_challenge(t, D, len(stir_points))
        if stir_points:
            cov = O.geometric_accumulate(D, cov, np.concatenate(srlc), np.concatenate(stir_points))
        for c, v in zip(srlc, stir_vals):
            the_sum = e_add(D, the_sum, e_mul(D, c, v))
        if check_invariants:
            assert np.array_equal(O.dot(D, vec, cov), the_sum), "dot(vector, covector) != the_sum"
        vec, cov, the_sum, fr = sumcheck_prove(t, rc.sumcheck, D, eng, vec, cov, the_sum)
        evaluation_point += fr
        prev_cfg, prev_ws = rc.irs, [new_w]
        vector_rlc = [e_one(D)]
    assert vec.size // D == cfg.final_sumcheck.initial_size
    for i in range(vec.size // D):
        t.prover_field(D, vec[i * D:(i + 1) * D])
    pow_prove(t, eng, cfg.final_pow)
    irs_open_many(t, prev_cfg, prev_ws)
    vec, cov, the_sum, fr = sumcheck_prove(t, cfg.final_sumcheck, D, eng, vec, cov, the_sum)
    evaluation_point += fr
    return evaluation_point, form_rlc


def verify(r, cfg: WhirConfig, evaluations: list, points: list):
    """whir::Config::receive_commitment + verify + FinalClaim::verify
    (src/protocols/whir/mod.rs:63-74, 95-105; verifier.rs:45-264), one commitment,
    MultilinearExtension linear forms."""
    return verify_batch(r, cfg, 1, [[e] for e in evaluations], points)


def verify_batch(r, cfg: WhirConfig, num_commitments: int, evaluations: list, points: list):
    """receive_commitment for each of num_commitments commitments (batch_size
    vectors each), then whir::Config::verify (verifier.rs:45-264) and
    FinalClaim::verify (whir/mod.rs:63-74) for MLE linear forms;
    evaluations[p][j] as in prove_batch."""
    D, eng = cfg.D, cfg.hash_engine
    B = cfg.initial.num_vectors
    V = num_commitments * B
    vassert(all(len(row) == V for row in evaluations), "evaluation matrix shape")
    commits = []
    for _ in range(num_commitments):
        root = r.prover_hash()
        opts = [r.verifier_field(D) for _ in range(cfg.initial.out_domain_samples)]
        ovals = [r.prover_field(D) for _ in range(cfg.initial.out_domain_samples * B)]
        commits.append((root, opts, ovals))
    oods_points, oods_matrix = [], []
    off = 0
    for root, opts, ovals in commits:
        for i, x in enumerate(opts):
            row = [ovals[i * B + j - off] if off <= j < off + B else r.prover_field(D) for j in range(V)]
            oods_points.append(x)
            oods_matrix.append(row)
        off += B
    vrlc = geometric_challenge(r, D, V)
    crlc = geometric_challenge(r, D, len(oods_points) + len(evaluations))
    form_rlc, oods_rlc = crlc[:len(evaluations)], crlc[len(evaluations):]
    the_sum = np.zeros(D, dtype=np.uint64)
    for c, row in zip(form_rlc, evaluations):
        the_sum = e_add(D, the_sum, e_mul(D, c, _row_dot(D, vrlc, np.concatenate(row), D)))
    for c, row in zip(oods_rlc, oods_matrix):
        the_sum = e_add(D, the_sum, e_mul(D, c, _row_dot(D, vrlc, np.concatenate(row), D)))
    round_constraints = [(oods_rlc, oods_points)]
    if not crlc:
        vassert(not the_sum.any())
        fr = [r.verifier_field(D) for _ in range(cfg.initial_sumcheck.num_rounds)]
        pow_verify(r, eng, cfg.initial_skip_pow)
    else:
        the_sum, fr = sumcheck_verify(r, cfg.initial_sumcheck, D, eng, the_sum)
    round_fr = [fr]
    prev = (cfg.initial, [c[0] for c in commits], 1, vrlc)
    for rc in cfg.rounds:
        nroot = r.prover_hash()
        npts = [r.verifier_field(D) for _ in range(rc.irs.out_domain_samples)]
        nvals = [r.prover_field(D) for _ in range(rc.irs.out_domain_samples)]
        pow_verify(r, eng, rc.pow)
        pts, rows = irs_verify_many(r, prev[0], eng, prev[1], prev[2])
        weights = _tensor(D, prev[3], list(eq_weights(D, round_fr[-1])))
        cw = list(npts) + [O.lift(D, [p]) for p in pts]
        vals = list(nvals) + [_row_dot(D, weights, row, prev[2]) for row in rows]
        srlc = geometric_challenge(r, D, len(vals))
        for c, v in zip(srlc, vals):
            the_sum = e_add(D, the_sum, e_mul(D, c, v))
        round_constraints.append((srlc, cw))
        the_sum, fr = sumcheck_verify(r, rc.sumcheck, D, eng, the_sum)
        round_fr.append(fr)
        prev = (rc.irs, [nroot], D, [e_one(D)])
    nf = cfg.final_sumcheck.initial_size
    final_vector = r.prover_field(D, nf)
    pow_verify(r, eng, cfg.final_pow)
    pts, rows = irs_verify_many(r, prev[0], eng, prev[1], prev[2])
    weights = _tensor(D, prev[3], list(eq_weights(D, round_fr[-1])))
    for p, row in zip(pts, rows):
        val = _row_dot(D, weights, row, prev[2])
        ev = O.univariate_evaluate(D, D, final_vector, O.lift(D, [p]))
        vassert(np.array_equal(ev, val), "final in-domain")
    the_sum, ffr = sumcheck_verify(r, cfg.final_sumcheck, D, eng, the_sum)
    round_fr.append(ffr)
    evaluation_point = [x for rf in round_fr for x in rf]
    poly_eval = O.multilinear_extend(D, final_vector, np.concatenate(ffr)) if ffr else final_vector[:D]
    # the_sum / poly_eval (verifier.rs:241-243): ark-ff panics on a zero divisor,
    # which happens for the zero vector; reported here as a rejection.
    vassert(np.asarray(poly_eval).any(), "final polynomial evaluates to zero (division by zero)")
    inv = _e_inv(D, poly_eval)
    linear_form_rlc = e_mul(D, the_sum, inv)
    nv_list = [cfg.initial.vector_size.bit_length() - 1] + [rc.irs.vector_size.bit_length() - 1 for rc in cfg.rounds]
    for ri, (coeffs, wts) in enumerate(round_constraints):
        nv = nv_list[ri]
        start = max(len(evaluation_point) - nv, 0)
        sub = evaluation_point[start:]
        for c, x in zip(coeffs, wts):
            linear_form_rlc = e_sub(D, linear_form_rlc, e_mul(D, c, univariate_mle(D, x, sub)))
    # FinalClaim::verify with MLE linear forms
    rlc = np.zeros(D, dtype=np.uint64)
    for c, pt in zip(form_rlc, points):
        if isinstance(pt, Covector):  # Covector::mle_evaluate, covector.rs:17-19
            val = O.multilinear_extend(D, np.asarray(pt.vector, dtype=np.uint64).reshape(-1),
                                       np.concatenate(evaluation_point))
        elif isinstance(pt, Univariate):  # univariate_evaluation.rs:46-56
            val = univariate_mle(D, np.asarray(pt.point, dtype=np.uint64), evaluation_point)
        else:
            val = eq_mle(D, pt, evaluation_point)
        rlc = e_add(D, rlc, e_mul(D, c, val))
    vassert(np.array_equal(rlc, linear_form_rlc), "final claim")
    r.check_eof()
    return evaluation_point


def univariate_mle(D, x, point):
    """UnivariateEvaluation::mle_evaluate, src/algebra/linear_form/univariate_evaluation.rs:46-56"""
    x2i = ext(D, x)
    res = e_one(D)
    for rr in reversed(point):
        res = e_mul(D, res, e_add(D, e_sub(D, e_one(D), rr), e_mul(D, rr, x2i)))
        x2i = e_mul(D, x2i, x2i)
    return res


def eq_mle(D, z, r):
    """MultilinearExtension::mle_evaluate: prod_i (z_i r_i + (1 - z_i)(1 - r_i))."""
    res = e_one(D)
    for zi, ri in zip(z, r):
        a = e_mul(D, zi, ri)
        b = e_mul(D, e_sub(D, e_one(D), zi), e_sub(D, e_one(D), ri))
        res = e_mul(D, res, e_add(D, a, b))
    return res


def _e_inv(D, a):
    # a^(p^D - 2)
    return O.pow_(D, ext(D, a), P**D - 2).reshape(-1) if P**D - 2 < 2**64 else _e_pow_big(D, a, P**D - 2)


def _e_pow_big(D, a, e):
    r = e_one(D)
    b = ext(D, a)
    while e:
        if e & 1:
            r = e_mul(D, r, b)
        b = e_mul(D, b, b)
        e >>= 1
    return r


def protocol_ids(cfg_desc: str, session: str = "whir_amd"):
    """Domain separator of this project's transcript: SHA3-512 / SHA3-256 of strings."""
    import hashlib
    return hashlib.sha3_512(cfg_desc.encode()).digest(), hashlib.sha3_256(session.encode()).digest()
