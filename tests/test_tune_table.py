"""Tuning-table text of the C ABI (sfxg_tune_import / sfxg_tune_export /
sfxg_tune_clear): host logic only, no device call. A line is
"shape|fe_threads|chan_wgs|xs_waves|mf_channels|mfp_wgs|mf_ks|chan_fp32|fe_fused";
trailing fields may be left out (at least five given) and read as 0."""
import pytest

from sfxg import capi


def test_tune_table_round_trip():
    capi.tune_clear()
    try:
        k8 = capi.tune_key(4096, 16, 0, False, True)
        k5 = capi.tune_key(1024, 24, 3, True, True)
        capi.tune_import(k8 + "|256|0|2|0|0|0|0|1\n" + k5 + "|128|0|0|0|0\n")
        lines = dict(ln.split("|", 1) for ln in capi.tune_export().strip().split("\n"))
        assert lines == {k8: "256|0|2|0|0|0|0|1", k5: "128|0|0|0|0|0|0|0"}
        # a later line for the same shape replaces the earlier one
        capi.tune_import(k8 + "|512|0|0|0|0|0|0|0\n")
        lines = dict(ln.split("|", 1) for ln in capi.tune_export().strip().split("\n"))
        assert lines[k8] == "512|0|0|0|0|0|0|0" and len(lines) == 2
    finally:
        capi.tune_clear()
    assert capi.tune_export() == ""


def test_tune_table_refuses_short_lines():
    capi.tune_clear()
    try:
        good = capi.tune_key(256, 4, 0, False, True) + "|64|0|0|0|0|0|0|1\n"
        with pytest.raises(capi.SfxgError):
            capi.tune_import(good + "nc=512 M=4|128|0|0\n")  # four fields
        assert capi.tune_export() == ""  # nothing of a refused text is added
        with pytest.raises(capi.SfxgError):
            capi.tune_import("no separator\n")
    finally:
        capi.tune_clear()
