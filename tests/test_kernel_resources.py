"""Register budgets of the step kernels, from the compiler's resource-usage
remarks that the build keeps (acoustic_isotropic_operators_amd/_build/
kernels_resource_usage.txt): no kernel spills registers except the 32 x 16
lockstep pair kernel's one register (measured faster at 6 waves per SIMD
than without the spill at 5) and the 21-offset imaging kernel's two
(faster at 4 waves than without them at 3), and the default kernels of the
Born benchmark keep the occupancy they were tuned at.  CPU only."""
import os
import re

import pytest

from acoustic_isotropic_operators_amd import build

ALLOWED_SPILLS = {
    "_ZN4aiso10k_fwd_pairILi32ELi16EEEvNS_8StepArgsE": 1,
    "_ZN4aiso14k_ext_img_tileILi10ELi7ELb0EEEvNS_12ExtBatchArgsEiiii": 2,
}
# MODE_SRC_WIN = 520, MODE_REC_WIN = 529, MODE_NL_FWD = 136
TUNED_OCCUPANCY = {
    "_ZN4aiso10k_fwd_pairILi32ELi16EEEvNS_8StepArgsE": 6,
    "_ZN4aiso8k_step_wILi32ELi32ELj520ELi6EEEvNS_8StepArgsE": 6,
    "_ZN4aiso8k_step_wILi32ELi32ELj529ELi6EEEvNS_8StepArgsE": 6,
    "_ZN4aiso10k_step_winILi32ELi32ELj136EEEvNS_11WinStepArgsE": 8,
    "_ZN4aiso14k_ext_img_tileILi10ELi7ELb0EEEvNS_12ExtBatchArgsEiiii": 4,
}


def _kernels():
    if not os.path.exists(build.RESOURCE_USAGE):
        pytest.skip("libaiso.so not built here (no resource-usage remarks)")
    rows, cur = {}, None
    for line in open(build.RESOURCE_USAGE):
        m = re.search(r"remark: Function Name: (\S+)", line)
        if m:
            cur = rows.setdefault(m.group(1), {})
            continue
        m = re.search(r"remark:\s+(VGPRs Spill|VGPRs|Occupancy \[waves/SIMD\]): (\d+)", line)
        if m and cur is not None:
            cur[m.group(1).split()[0] + (" Spill" if "Spill" in m.group(1) else "")] = int(m.group(2))
    assert rows, "no kernel remarks in the build output"
    return rows


def test_no_unexpected_register_spills():
    spills = {k: v.get("VGPRs Spill", 0) for k, v in _kernels().items() if v.get("VGPRs Spill", 0)}
    assert spills == ALLOWED_SPILLS


def test_default_kernels_keep_their_tuned_occupancy():
    rows = _kernels()
    for name, occ in TUNED_OCCUPANCY.items():
        assert name in rows, name
        assert rows[name]["Occupancy"] == occ, (name, rows[name])
