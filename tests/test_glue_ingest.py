"""Ingest of the reference's glue subdomains (bfamx_mortars_from_glue):
the reference-format arrays EToEm/EToFm/EToHm/EToEp/EToHp/EToOp, built here
by the oracle's restatement of the reference's face mapping and glue
construction (bfam.c:4666-4796, 7480-7492) from the element geometry alone,
must give exactly the mortar set the library's own octree builder makes.
The reference's mesh probe (a 4x4x4 brick with one tree refined) has 48
hanging glue elements, which become 24 mortars.  CPU only."""
import json
import os

import numpy as np
import pytest

import bfam_amd as bx
import bfam_amd.workloads as W
from oracle.glue import glue_subdomains

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def mesh_arrays(cfg):
    """element boxes in mesh order with subdomain and local index"""
    K = cfg["octree"]["K"]
    x0, h = np.zeros((K, 3)), np.zeros((K, 3))
    sub_of, local_of = np.zeros(K, np.int64), np.zeros(K, np.int64)
    for si, sm in enumerate(cfg["subs"]):
        x0[sm.gid], h[sm.gid] = sm.x0, sm.h
        sub_of[sm.gid] = si
        local_of[sm.gid] = np.arange(sm.K)
    return x0, h, sub_of, local_of


def as_set(m):
    return {tuple(int(r[k]) for k in ("s_sub", "s_elem", "s_face", "b_sub", "b_elem",
                                      "b_face", "quad")) for r in m}


@pytest.mark.parametrize("orders", [(5, 4, 3), (4, 4, 4), (3, 5, 7)])
def test_ingested_glue_gives_the_builders_mortars(orders):
    cfg = W.elastic3d_octree(n_base=4, levels=2, radius=0.2, orders=orders, seed=5,
                             material=(1.3, 2.0, 1.1))
    glues = glue_subdomains(*mesh_arrays(cfg))
    # every piece appears once from each side
    assert sum(g["K"] for g in glues) == 2 * len(cfg["mortars"])
    got = bx.mortars_from_glue(glues)
    assert len(got) == len(cfg["mortars"])
    assert as_set(got) == as_set(cfg["mortars"])


def test_reference_mesh_probe_glue_count():
    """the probe mesh: one subdomain, one local glue of 48 hanging pieces"""
    probe = json.load(open(os.path.join(GOLDEN, "reference_constants.json")))["mesh_probe"]
    cfg = W.elastic3d_octree(n_base=probe["brick"][0], levels=1, radius=0.01,
                             orders=(probe["N"], probe["N"]), center=(0.375, 0.625, 0.125),
                             material=(1.0, 2.0, 1.0))
    glues = glue_subdomains(*mesh_arrays(cfg))
    assert len(glues) == 1 and glues[0]["K"] == probe["hanging_glue_elements"]
    assert cfg["octree"]["K"] == probe["volume_elements"]
    m = bx.mortars_from_glue(glues)
    assert len(m) == 24 and as_set(m) == as_set(cfg["mortars"])


def test_inconsistent_or_rotated_glue_is_refused():
    cfg = W.elastic3d_octree(n_base=4, levels=1, radius=0.2, orders=(5, 3), seed=5,
                             material=(1.3, 2.0, 1.1))
    glues = glue_subdomains(*mesh_arrays(cfg))
    bad = [dict(g) for g in glues]
    bad[0]["EToEp"] = np.roll(bad[0]["EToEp"], 1)  # partners no longer name each other
    with pytest.raises(bx.BfamxError) as e:
        bx.mortars_from_glue(bad)
    assert e.value.code == bx.EINVAL
    rot = [dict(g) for g in glues]
    rot[0]["EToOp"] = rot[0]["EToOp"].copy()
    rot[0]["EToOp"][0] = 3
    with pytest.raises(bx.BfamxError) as e:
        bx.mortars_from_glue(rot)
    assert e.value.code == bx.EUNSUPPORTED
    lone = [g for g in glues if g["sub_m"] != glues[0]["sub_p"] or g["sub_p"] != glues[0]["sub_m"]]
    if len(lone) < len(glues):
        with pytest.raises(bx.BfamxError) as e:
            bx.mortars_from_glue(lone)  # a glue without its partner
        assert e.value.code == bx.EINVAL
