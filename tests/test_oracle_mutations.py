"""Each pin in tests/oracle_pins.py must reject a plausible mistake: copies of
the oracle are built with one deliberate one-line mistake planted
(OR_MUTATION k, see oracle/or_internal.h), and the check that pins that
rule must fail on the copy.  CPU only."""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import oracle_py
import oracle_pins as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# mutation -> (what is planted, the check that must reject it)
MUTATIONS = {
    1: ("inner search window: per-axis min and max extent swapped",
        lambda o: P.check_volume_search(o, "hex", n=5, seed=7, jitter=0.2, min_clean=0.75)),
    2: ("flat 2D search: A wins an exact A/B tie", P.check_flat_tie),
    3: ("hex metric: the principal-axes sub-tet dropped", P.check_hex_metric),
    4: ("projection: blend weight squared", P.check_projection),
    5: ("flat 2D search: written whenever any level improved", P.check_flat_write_rule),
    6: ("inner search: the highest candidate index wins a tie", P.check_tet_tie),
    7: ("surface 2D window: tet and hex distance rules swapped", lambda o: P.check_surface_search(o, "hex", 2)),
    8: ("surface 2D search: B wins an exact A/B tie", P.check_surface_tie),
    9: ("nearest-neighbour walk: feature-class filter dropped", P.check_nn_walk),
    10: ("feature 1D window: tet and hex distance rules swapped", lambda o: P.check_surface_search(o, "tet", 1)),
    11: ("projection: a feature neighbour always uses its first sector", P.check_projection),
    12: ("learned window: a tie at a positive q(v) falls back", lambda o: P.check_learned_window(o, "hex", 1, 2.0)),
    13: ("learned window: halved per level instead of 2/3", lambda o: P.check_learned_window(o, "tet", 2, 0.5)),
    14: ("surface-mesh fan window from the nearest fan vertex", lambda o: P.check_fan_search(o, "tri", 2, "sphere")),
    15: ("flat boundary 1D pass: the open fan closed", P.check_flat1d_search),
    16: ("feature sectors: a cut mapped to the sector it opens", P.check_sectors),
    17: ("2-ring fit: neighbours shared by two one-ring vertices taken twice", P.check_ring2_fits),
    18: ("inner hex search: the nine sub-tets scored under the eight-corner metric", P.check_hex_corner8_rings),
}


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    d = tmp_path_factory.mktemp("mutants")

    def build(k):
        out = str(d / f"liboracle_m{k}.so")
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "mutant", f"MUTATION={k}", f"OUT={out}"],
                       check=True)
        return k, out

    with ThreadPoolExecutor(max_workers=4) as ex:
        return dict(ex.map(build, MUTATIONS))


@pytest.mark.parametrize("k", sorted(MUTATIONS))
def test_planted_mistake_is_rejected(oracle, mutants, k):
    what, check = MUTATIONS[k]
    check(oracle)                                   # the oracle itself passes
    with oracle_py.use_library(mutants[k]):
        with pytest.raises(AssertionError):
            check(oracle_py)
