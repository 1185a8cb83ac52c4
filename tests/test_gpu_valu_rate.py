"""The integer VALU issue-rate microbenchmark (slime_valu_rate) that the
scan's roofline is read against: every instruction kind runs, no kind
exceeds MI355X's published 32-bit lane rate (157.3 TFLOP/s FP32 = 78.6 T
lane-ops/s; a 2% margin for timer granularity), and single-cycle VOP2 forms
issue faster than the VOP3 forms."""
import pytest

pytestmark = pytest.mark.gpu

PEAK_T = 157.3 / 2


def test_valu_issue_rates(gpu):
    import slime_amd
    rates = {}
    for kind, name in enumerate(slime_amd.VALU_KINDS):
        r = slime_amd.valu_rate(kind, 16384)
        assert r is not None, name
        rates[name] = r / 1e12
    print("VALU issue rates (T lane-instructions/s):", rates, flush=True)
    for name, t in rates.items():
        assert 1.0 < t <= PEAK_T * 1.02, (name, t)
    assert min(rates["v_add_u32"], rates["v_xor_b32"]) > max(rates["v_mul_lo_u32"],
                                                               rates["v_mad_u32_u24"])
    assert slime_amd.lib().slime_valu_rate(10, 16, None, None) == -1
    assert slime_amd.lib().slime_valu_rate(0, 0, None, None) == -1
