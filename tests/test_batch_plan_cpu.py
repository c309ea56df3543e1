"""The serving planner (retrocam_amd.batch.plan_requests): shortest job first,
requests too small to fill the GPU packed into shared launches."""
import random

from retrocam_amd.batch import FILL_UNITS, ViewRequest, plan_requests, request_units
from retrocam_amd.scene import CameraSpec, PtParams


def req(w, h, spp):
    cam = CameraSpec((0, 0, 0), (0, 0, -1), (1, 0, 0), (0, 1, 0), w, h)
    return ViewRequest(0, cam, PtParams(1, 1.0, 6), spp)


def test_plan_is_shortest_first_and_covers_every_request_once():
    rng = random.Random(3)
    for _ in range(50):
        reqs = [req(rng.choice([64, 320, 640, 1280]), rng.choice([36, 180, 360, 720]), rng.randint(1, 128))
                for _ in range(rng.randint(1, 40))]
        groups = plan_requests(reqs)
        flat = [i for g in groups for i in g]
        assert sorted(flat) == list(range(len(reqs)))
        units = [request_units(reqs[i]) for i in flat]
        assert units == sorted(units)                      # ascending work
        for g in groups:
            total = sum(request_units(reqs[i]) for i in g)
            assert len(g) == 1 or total <= FILL_UNITS      # only small requests share a launch
        for a, b in zip(groups, groups[1:]):               # a group closes only when the next request
            nxt = request_units(reqs[b[0]])                # would overflow it
            assert sum(request_units(reqs[i]) for i in a) + nxt > FILL_UNITS


def test_plan_ties_keep_arrival_order_and_respects_the_view_limit():
    reqs = [req(64, 36, 4) for _ in range(10)]
    assert plan_requests(reqs) == [list(range(10))]
    assert plan_requests(reqs, max_views=4) == [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9]]
    big = req(1280, 720, 128)
    assert request_units(big) > FILL_UNITS
    assert plan_requests([big, req(64, 36, 4), big]) == [[1], [0], [2]]
