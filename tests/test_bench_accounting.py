"""bench.py's roofline accounting: the per-unit algorithmic bytes its docstring states, term by term,
and the committed profile summaries it reports beside them (HBM traffic, occupancy)."""
import bench

ZERO = {k: 0 for k in ("closest_rays", "closest_nodes", "closest_tris", "shadow_rays", "shadow_nodes", "shadow_tris",
                       "closest_hits")}


def bytes_for(pixels=0, delta_lights=0, env=False, **counts):
    c = dict(ZERO)
    c.update(counts)
    return bench.algorithmic_bytes(c, pixels, delta_lights, env)


def test_nothing_traced_moves_only_the_film():
    b = bytes_for(pixels=10)
    assert b == {"wf_extend": 0, "wf_shade": 48 * 10, "wf_shadow": 0, "wf_trace": 0}


def test_a_node_and_a_triangle_cost_one_line_and_three_words():
    assert bytes_for(closest_nodes=1)["wf_extend"] == 128
    assert bytes_for(closest_tris=1)["wf_extend"] == 48
    assert bytes_for(shadow_nodes=1)["wf_shadow"] == 128
    assert bytes_for(shadow_tris=1)["wf_shadow"] == 48
    # the merged launch carries both kinds of traversal and nothing of the resolve step
    assert bytes_for(closest_nodes=2, shadow_nodes=3)["wf_trace"] == 5 * 128


def test_a_camera_ray_is_regenerated_and_a_later_ray_is_read():
    # one pixel, one closest ray: the depth-0 ray costs its 16-byte hit only
    assert bytes_for(pixels=1, closest_rays=1)["wf_extend"] == 16
    # a second ray of that pixel (depth 1) is read from the state: 24 B more
    assert bytes_for(pixels=1, closest_rays=2)["wf_extend"] == 2 * 16 + 24
    # and its path state is read and written by the shading (17 words each way)
    s1 = bytes_for(pixels=1, closest_rays=1)["wf_shade"]
    s2 = bytes_for(pixels=1, closest_rays=2)["wf_shade"]
    assert s2 - s1 == 16 + 2 * 68


def test_a_hit_reads_the_shading_tables_and_the_lights():
    base = bytes_for(closest_hits=1)["wf_shade"]
    assert base == 244 + 24
    assert bytes_for(closest_hits=1, delta_lights=2)["wf_shade"] - base == 2 * 64
    assert bytes_for(closest_hits=1, env=True)["wf_shade"] - base == 48


def test_a_shadow_ray_costs_its_record_query_and_visibility():
    b = bytes_for(shadow_rays=1)
    assert b["wf_shade"] == 44
    assert b["wf_trace"] == 4 + 28 + 4
    assert b["wf_shadow"] == 4 + 28 + 4 + 16


def test_committed_summaries_cover_the_dominant_kernel():
    traffic = bench.traffic_from_profiles("wf_trace")
    assert traffic is not None and 10e6 < traffic < 1e9
    occ = bench.occupancy_from_profiles("wf_trace")
    assert occ is not None
    assert len(occ["by_depth"]) == 9  # depths 0-7 and the last depth's shadow rays
    assert 0 < occ["time_weighted"] <= max(occ["by_depth"]) <= occ["limit_waves_per_cu"] <= 32
    shade = bench.occupancy_from_profiles("wf_shade")
    assert shade is not None and len(shade["by_depth"]) == 8
    assert bench.occupancy_from_profiles("no_such_kernel") is None
