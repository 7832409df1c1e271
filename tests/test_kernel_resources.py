"""Register budget of the measured kernels (CPU; reads what the build wrote).

The churn kernels call the allocator's out-of-line paths, and a kernel's
register count is the largest of its call graph's.  The batched churn sits
at exactly 128 VGPRs, the most that still gives 4 waves per SIMD on gfx950:
a change to a rarely taken path (block requests, refills) that raised it
to 132-136 VGPRs cut occupancy to 3 and the headline from 4.05 to 3.51 G
pairs/s, with nothing failing.  This pins the occupancy and the scratch
size of the kernels the bench reports.
"""
import json
import os

import pytest

from gallatin_amd import build as B

BUDGET = {
    # kernel (mangled-name prefix): (minimum waves per SIMD, maximum scratch bytes per lane)
    "_Z7k_churn": (4, 32),
    "_Z15k_churn_batchedILi1024E": (4, 32),
    "_Z7k_plain": (4, 32),
    "_ZN12_GLOBAL__N_18k_streamILi1024ELi1EE": (4, 32),
    "_ZN12_GLOBAL__N_18k_streamILi256ELi0EE": (4, 32),
}


def _resources():
    if not os.path.exists(B.RESOURCES):
        pytest.skip("the HIP library has not been built here (run build())")
    with open(B.RESOURCES) as f:
        return json.load(f)


def test_parse_resource_usage_reads_compiler_remarks():
    text = "\n".join([
        "x.hip:1:1: remark: Function Name: _Z3fooPi [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark:     TotalSGPRs: 12 [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark:     VGPRs: 7 [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark:     AGPRs: 0 [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark:     ScratchSize [bytes/lane]: 16 [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark:     Occupancy [waves/SIMD]: 8 [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:9:1: remark: Function Name: _Z3barv [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:9:1: remark:     VGPRs: 130 [-Rpass-analysis=kernel-resource-usage]",
    ])
    got = B.parse_resource_usage(text)
    assert got == {"_Z3fooPi": {"sgprs": 12, "vgprs": 7, "agprs": 0, "scratch": 16,
                                "occupancy": 8},
                   "_Z3barv": {"vgprs": 130}}


@pytest.mark.parametrize("prefix", sorted(BUDGET))
def test_measured_kernels_keep_their_occupancy(prefix):
    res = _resources()
    hits = {k: v for k, v in res.items() if k.startswith(prefix)}
    assert hits, f"no kernel {prefix}* in the build's resource table"
    min_occ, max_scratch = BUDGET[prefix]
    for name, r in hits.items():
        assert r["occupancy"] >= min_occ, (name, r)
        assert r["scratch"] <= max_scratch, (name, r)
