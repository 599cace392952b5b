"""The oracle's pin tests must catch planted mistakes.

oracle/mutants.py builds copies of the oracle, each with one plausible one-line mistake in a
different function (bonded, bending, contact, friction, CenterOfMass, volumes, grid, division,
integration, Philox, the polar method's logarithm, the system-centre correction), and runs
tests/test_oracle_pins.py on each.  Every mutant must make a pin test fail, and the unchanged
source built the same way must pass."""
import os
import shutil
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ to build the mutants")
def test_every_planted_mistake_fails_a_pin_test():
    from oracle import mutants
    res = mutants.check(workers=4)
    assert res.pop("control").startswith("passed"), "the unchanged oracle fails its pins"
    survivors = {k: v for k, v in res.items() if not v.startswith("killed")}
    assert not survivors, survivors
    assert len(res) == len(mutants.MUTANTS) >= 20
