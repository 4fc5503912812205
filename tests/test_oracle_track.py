"""Oracle track management against TrackInit.test.cc golden arrays (CPU).

Golden data: tests/golden/track_init.json from
test/celeritas/track/TrackInit.test.cc:228-500.
"""

import json
import os

import numpy as np
import pytest

from tests import oracle_bind as ob
from tests import problems
from celeritas_amd import make_primaries

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "track_init.json")))


def as_int(v):
    return [-1 if x == 0xFFFFFFFF else int(x) for x in v]


def primaries(n):
    # TrackInitTestBase::make_primaries: E = 1 + i MeV at origin along +z
    return make_primaries(
        {"particle_id": 0, "energy": 1.0 + i, "pos": [0, 0, 0], "dir": [0, 0, 1]}
        for i in range(n))


def init_tracks(st):
    st.action("initialize-tracks")
    st.action("pre-step")


@pytest.fixture(scope="module")
def problem():
    return ob.Problem(problems.two_boxes_track_init())


def test_run(problem):
    g = GOLDEN["run"]
    st = ob.State(problem, g["num_tracks"])
    assert as_int(st.get_u32("vacancies")) == g["initial_vacancies"]
    st.insert_primaries(primaries(g["num_primaries"]))
    st.action("extend-from-primaries")
    assert as_int(st.get_u32("init_ids")) == g["primary_init_ids"]
    init_tracks(st)
    assert as_int(st.get_u32("track_id")) == g["init_track_ids"]
    assert as_int(st.get_u32("parent_id")) == g["init_parent_ids"]
    st.mock_interact(g["alloc"], g["alive"])
    st.action("extend-from-secondaries")
    assert as_int(st.get_u32("vacancies")) == g["vacancies_after"]
    init_ids = as_int(st.get_u32("init_ids"))
    assert len(init_ids) == g["num_inits_after"]
    assert init_ids[:2] == g["first_two_init_ids"]
    assert all(0 <= i <= g["init_id_max"] for i in init_ids)
    init_tracks(st)
    tids = as_int(st.get_u32("track_id"))
    assert all(0 <= i <= g["init_id_max"] for i in tids)
    for slot, tid in g["survivor_track_ids"].items():
        assert tids[int(slot)] == tid
    assert sorted(as_int(st.get_u32("parent_id"))) == g["sorted_parent_ids_after_init"]


def test_primaries(problem):
    g = GOLDEN["primaries"]
    n = g["num_tracks"]
    st = ob.State(problem, n)
    init_tracks(st)
    alive = [i % 2 for i in range(n)]
    for i in range(g["num_sets"]):
        st.insert_primaries(primaries(g["num_primaries"]))
        st.action("extend-from-primaries")
        init_tracks(st)
        st.mock_interact([0] * n, alive)
        st.action("extend-from-secondaries")
        c = st.counters()
        assert c["num_initializers"] == i * n // 2
        assert c["num_vacancies"] == n // 2
    assert as_int(st.get_u32("track_id")) == g["expected_track_ids"]
    assert as_int(st.get_u32("parent_id")) == g["expected_parent_ids"]
    assert as_int(st.get_u32("vacancies")) == g["expected_vacancies"]
    assert as_int(st.get_u32("init_ids")) == g["expected_init_ids"]


def test_extend_from_secondaries(problem):
    g = GOLDEN["extend_from_secondaries"]
    n = g["num_tracks"]
    st = ob.State(problem, n)
    st.insert_primaries(primaries(g["num_primaries"]))
    st.action("extend-from-primaries")
    for i in range(4):
        init_tracks(st)
        st.mock_interact(g["alloc"], g["alive"])
        st.action("extend-from-secondaries")
        assert as_int(st.get_u32("vacancies")) == g["expected_vacancies"]
        ids = as_int(st.get_u32("init_ids"))
        assert len(ids) == g["num_init_ids"]
        assert all(8 + i * 4 <= x <= 11 + i * 4 for x in ids)
        tids = as_int(st.get_u32("track_id"))
        assert all(0 <= x < n + (i + 1) * 4 for x in tids)
        pids = as_int(st.get_u32("parent_id"))
        for slot in range(n):
            if g["alive"][slot]:
                assert pids[slot] == -1


def test_add_more_primaries(problem):
    g = GOLDEN["add_more_primaries"]
    st = ob.State(problem, g["num_tracks"])
    for size, expected in zip(g["sizes"], g["expected_num_inits"]):
        st.insert_primaries(primaries(size))
        st.action("extend-from-primaries")
        assert st.counters()["num_initializers"] == expected


def test_double_insert_rejected(problem):
    st = ob.State(problem, 8)
    st.insert_primaries(primaries(2))
    with pytest.raises(ob.OracleError):
        st.insert_primaries(primaries(4))
