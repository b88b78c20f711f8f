"""The committed golden fixtures (tests/golden/, written by make_golden.py):
the oracle reproduces them on the CPU, and the GPU engine reproduces them
bit for bit. The six 1080p packet sizes are the README's own
(README.md:173-180); the hashes pin the bytes and samples around them.
"""

import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
README_SIZES = [187_677, 425_420, 613_107, 917_564, 1_376_337, 2_085_373]
PROFILES = [(0, 2, 10), (1, 2, 10), (2, 2, 10), (3, 2, 10), (4, 3, 12), (5, 3, 12)]


def sha(b):
    return hashlib.sha256(bytes(b)).hexdigest()


def readme():
    return json.load(open(os.path.join(HERE, "readme_1080p.json")))


def seeds():
    out = []
    for e in json.load(open(os.path.join(HERE, "seeds.json"))):
        out.append((open(os.path.join(HERE, e["file"]), "rb").read(), e))
    return out


def test_fixture_sizes_are_the_readmes():
    assert [e["size"] for e in readme()] == README_SIZES


@pytest.mark.slow
@pytest.mark.parametrize("i", range(6))
def test_oracle_reproduces_the_readme_fixtures(i):
    prof, chroma, depth = PROFILES[i]
    e = readme()[i]
    pkt = O.encode(O.test_frame(1920, 1080, chroma, depth, 24, 7), O.config(prof))
    assert (len(pkt), sha(pkt)) == (e["size"], e["packet_sha256"])
    assert sha(O.decode(pkt).data.astype("<u2").tobytes()) == e["samples_sha256"]


def test_oracle_reproduces_the_seed_fixtures():
    for pkt, e in seeds():
        assert len(pkt) == e["size"]
        d = O.decode(pkt)
        assert sha(d.data.astype("<u2").tobytes()) == e["samples_sha256"]
        assert (None if d.alpha is None else sha(d.alpha.astype("<u2").tobytes())) == e["alpha_sha256"]


@pytest.mark.gpu
@pytest.mark.parametrize("i", range(6))
def test_gpu_reproduces_the_readme_fixtures(i):
    import prores_amd as P
    prof, chroma, depth = PROFILES[i]
    e = readme()[i]
    f = O.test_frame(1920, 1080, chroma, depth, 24, 7)
    pf = P.Frame(f.width, f.height, f.chroma, f.bit_depth, f.data.copy(), None, f.interlace,
                 P.Metadata(*f.metadata))
    print("case: golden readme", prof, flush=True)
    pkt = P.Encoder(P.Config(prof)).encode(pf)
    assert (len(pkt), sha(pkt)) == (e["size"], e["packet_sha256"])
    assert sha(P.Decoder().decode(pkt).data.astype("<u2").tobytes()) == e["samples_sha256"]


@pytest.mark.gpu
def test_gpu_reproduces_the_seed_fixtures():
    import prores_amd as P
    dec = P.Decoder()
    for pkt, e in seeds():
        print("case: golden seed", e["file"], flush=True)
        d = dec.decode(pkt)
        assert sha(d.data.astype("<u2").tobytes()) == e["samples_sha256"]
        assert (None if d.alpha is None else sha(d.alpha.astype("<u2").tobytes())) == e["alpha_sha256"]
