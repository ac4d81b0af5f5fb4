"""The product's parameter selection (C++, whir_config_new) against the oracle's
restatement of src/protocols/whir/config.rs:16-152 (host only, no GPU)."""
import itertools

import pytest

from oracle import oracle as O
from oracle import whir_protocol as WP
from whir_amd import prover as WPROV

CASES = list(itertools.product(
    [1, 2, 3],            # degree
    [10, 16, 20, 24, 26],  # num_variables
    [(4, 4), (1, 1), (2, 3), (3, 2)],  # folding factors
    [1, 2, 3],            # starting log inverse rate
    [False, True],        # unique decoding
    [(100, 20), (128, 20), (32, 0), (80, 10)],  # security, pow bits
))


def oracle_desc(D, nv, k0, k, lir, ud, sec, pb, engine):
    p = WP.ProtocolParameters(security_level=sec, pow_bits=pb, initial_folding_factor=k0, folding_factor=k,
                              unique_decoding=ud, starting_log_inv_rate=lir, hash_engine=engine)
    return WP.WhirConfig.new(1 << nv, p, D)


@pytest.mark.parametrize("D,nv,ff,lir,ud,sp", CASES[::7] + [(2, 24, (4, 4), 2, False, (100, 20)),
                                                            (2, 20, (4, 4), 1, False, (100, 20)),
                                                            (3, 26, (4, 4), 1, False, (100, 20))])
def test_config_matches_oracle(D, nv, ff, lir, ud, sp):
    k0, k = ff
    sec, pb = sp
    if D == 1 and not ud and sec > 100:
        pytest.skip("list decoding over the 64-bit field needs more OOD bits than it has")
    engine = O.ENGINE_BLAKE3
    try:
        ref = oracle_desc(D, nv, k0, k, lir, ud, sec, pb, engine)
    except (AssertionError, ValueError):
        with pytest.raises(Exception):
            WPROV.Config(1 << nv, WPROV.ProtocolParameters(degree=D, security_level=sec, pow_bits=pb,
                                                           initial_folding_factor=k0, folding_factor=k,
                                                           unique_decoding=ud, starting_log_inv_rate=lir,
                                                           hash_engine=engine))
        return
    cfg = WPROV.Config(1 << nv, WPROV.ProtocolParameters(degree=D, security_level=sec, pow_bits=pb,
                                                         initial_folding_factor=k0, folding_factor=k,
                                                         unique_decoding=ud, starting_log_inv_rate=lir,
                                                         hash_engine=engine))
    assert cfg.describe() == ref.describe()
    cfg.disable_pow()
    ref.disable_pow()
    assert cfg.describe() == ref.describe()


@pytest.mark.parametrize("B", [2, 3, 8])
@pytest.mark.parametrize("D,nv", [(2, 16), (3, 20)])
def test_config_batch_size_matches_oracle(B, D, nv):
    """batch_size only widens the initial commitment (irs_commit.rs:181-190)."""
    p = WP.ProtocolParameters(security_level=100, pow_bits=20, starting_log_inv_rate=1, batch_size=B)
    ref = WP.WhirConfig.new(1 << nv, p, D)
    cfg = WPROV.Config(1 << nv, WPROV.ProtocolParameters(degree=D, security_level=100, pow_bits=20,
                                                         starting_log_inv_rate=1, batch_size=B))
    assert cfg.describe() == ref.describe()
    assert cfg.describe()["initial"]["num_vectors"] == B
