"""Occupancy of the hot kernels, from the compiler's resource report of the
last library build (ntlp_amd/kernel_resources.json, written by build()).

The mixed-precision BE step is latency-bound: at 127 VGPRs it runs 4 waves per
SIMD and takes 1.30 ms at 20M droplets; one register more drops it to 3 waves
and 1.44 ms.  The LDS box scatter runs 6 workgroups per CU (LDS-limited) as
long as it needs at most 80 VGPRs.  Neither may spill to scratch."""
from __future__ import annotations

import json
import os

import pytest

from ntlp_amd import build

STEP_MIXED_BE = ("_ZN4ntlp6k_stepILi2ELi1ELi1ELb0ELb1EEEvNS_8StepArgsE",   # library step (multi-step calls)
                 "_ZN4ntlp6k_stepILi2ELi1ELi1ELb1ELb0EEEvNS_8StepArgsE")   # serving step
SCATTER_BOX = ("_ZN4ntlp13k_scatter_boxILi8ELi8ELi8EdEEvNS_8TileArgsE",   # FP64 amounts
               "_ZN4ntlp13k_scatter_boxILi8ELi8ELi8EfEEvNS_8TileArgsE")   # FP32 amounts (mixed precision)


@pytest.fixture(scope="module")
def resources():
    if not os.path.exists(build.RESOURCES):
        pytest.skip("library not built here (no resource report)")
    with open(build.RESOURCES) as f:
        return json.load(f)


def test_report_parses_the_compiler_remarks():
    text = ("k.hpp:1:1: remark: Function Name: _Zfoo [-Rpass-analysis=kernel-resource-usage]\n"
            "k.hpp:1:1: remark:     VGPRs: 127 [-Rpass-analysis=kernel-resource-usage]\n"
            "k.hpp:1:1: remark:     ScratchSize [bytes/lane]: 0 [-Rpass-analysis=kernel-resource-usage]\n"
            "k.hpp:1:1: remark:     Occupancy [waves/SIMD]: 4 [-Rpass-analysis=kernel-resource-usage]\n")
    assert build.parse_resources(text) == {"_Zfoo": {"VGPRs": 127, "ScratchSize": 0, "Occupancy": 4}}


@pytest.mark.parametrize("name", STEP_MIXED_BE)
def test_mixed_be_step_keeps_four_waves_per_simd(resources, name):
    r = resources[name]
    assert r["VGPRs"] <= 128 and r["Occupancy"] >= 4, r
    assert r["ScratchSize"] == 0 and r["VGPRs Spill"] == 0, r


@pytest.mark.parametrize("name", SCATTER_BOX)
def test_scatter_box_fits_six_workgroups_per_cu(resources, name):
    r = resources[name]
    assert r["VGPRs"] <= 80, r
    assert r["ScratchSize"] == 0 and r["VGPRs Spill"] == 0, r
