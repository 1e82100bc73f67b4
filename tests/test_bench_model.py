"""CPU checks of bench.py's algorithmic-byte model, the numerator of the
roofline's `achieved`: per-launch bytes of each plan step at the benchmark
shape (n = 2^16, L = 24, K = 6, dnum = 4), written out here term by term."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

N, L, K, ALPHA, NP, B = 1 << 16, 24, 6, 6, 4, 64
T = 8 * N                                   # one tower
FUSED_KSI = ["tensor", "ntt_inv", "conv", "ntt_fwd_a", "ksi", "ksi", "ntt_inv", "rsv", "ntt_fwd"]
FUSED_MODUP = ["tensor", "ntt_inv_b", "ntt_inv", "modup", "ksi", "ksi", "ntt_inv", "rsv", "ntt_fwd"]
FUSED_MODDOWN = ["tensor", "ntt_inv_b", "ntt_inv", "modup", "ksi", "ksi", "ntt_inv", "ntt_inv_b", "moddown",
                 "ntt_fwd_b"]


def test_fused_plan_steps():
    steps = bench.step_bytes(N, L, K, ALPHA, NP, FUSED_KSI, B)
    names = [s[0] for s in steps]
    assert names == ["tensor", "intt_modup", "conv_modup", "ntt_digits_col", "rowpass_inner_q",
                     "rowpass_inner_p", "intt_moddown", "moddown_rescale_prologue", "ntt_rescale_epilogue"]
    by = {s[0]: s[1] for s in steps}
    # each part converts its 6 own towers to the other 18 Q towers and the 6 P towers
    assert by["conv_modup"] == T * (4 * 6 + 4 * 24)
    # column pass of the 96 digit towers: half of each transform's 16 n B
    assert by["ntt_digits_col"] == T * 96
    # Q rows: 3 transformed digits (their row-pass half), the own digit, 8 key
    # rows once per batch, two outputs -- per op, 24 rows
    assert by["rowpass_inner_q"] == pytest.approx(T * 24 * (3 + 1 + 8 / B + 2))
    # P rows: 4 transformed digits, 8 key rows per batch, two outputs, 6 rows
    assert by["rowpass_inner_p"] == pytest.approx(T * 6 * (4 + 8 / B + 2))
    # per launch at B = 64, Q rows: 24 x (6 x 64 + 8) towers
    assert round(by["rowpass_inner_q"] * B) == 24 * (6 * 64 + 8) * T
    # prologue with the dropped tower: per component 6 sources + dt + dr in, 23 towers out
    assert by["moddown_rescale_prologue"] == T * 2 * (6 + 2 + 23)


def test_key_rows_count_once_per_batch():
    one = {s[0]: s[1] for s in bench.step_bytes(N, L, K, ALPHA, NP, FUSED_KSI, 1)}
    many = {s[0]: s[1] for s in bench.step_bytes(N, L, K, ALPHA, NP, FUSED_KSI, 64)}
    # the whole difference is the key rows: 8 towers per row, shared by the batch
    assert one["rowpass_inner_q"] - many["rowpass_inner_q"] == pytest.approx(T * 24 * 8 * (1 - 1 / 64))
    for k in one:
        if not k.startswith("rowpass"):
            assert one[k] == many[k], k


def test_kernel_split_of_two_pass_transforms():
    # 16 n B per tower split evenly; epilogue operands go to the row pass
    assert bench.kernel_bytes("x", "ntt_fwd", 2, 1000, 0) == [500, 500]
    assert bench.kernel_bytes("x", "ntt_fwd", 2, 1000, 400) == [300, 700]
    assert bench.kernel_bytes("x", "ksi", 1, 777, 0) == [777]


def test_unfused_plan_has_the_reference_stage_order():
    kinds = ["tensor", "ntt_inv", "conv", "ntt_fwd", "inner", "ntt_inv", "conv", "ntt_fwd", "combine",
             "ntt_inv", "switch", "ntt_fwd", "fuse"]
    names = [s[0] for s in bench.step_bytes(N, L, K, ALPHA, NP, kinds, B)]
    assert names[3:5] == ["ntt_digits", "key_inner_product"] and names[-1] == "rescale_fuse"


def test_fused_modup_plan_steps():
    """With the fused ModUp the same algorithmic bytes are charged as the
    unfused stages they replace: the inverse transforms (24 towers, row pass
    here, column pass in the fused kernel), the conversion and the digits'
    column pass; the extra inverse transforms of t0/t1's last tower stand
    alone."""
    fused = {s[0]: s[1] for s in bench.step_bytes(N, L, K, ALPHA, NP, FUSED_MODUP, B)}
    apart = {s[0]: s[1] for s in bench.step_bytes(N, L, K, ALPHA, NP, FUSED_KSI, B)}
    assert list(fused) == ["tensor", "intt_modup_rowpass", "intt_extra", "modup_fused", "rowpass_inner_q",
                           "rowpass_inner_p", "intt_moddown", "moddown_rescale_prologue", "ntt_rescale_epilogue"]
    assert fused["intt_modup_rowpass"] == T * 24
    assert fused["intt_extra"] == 2 * T * 2
    # INTT column passes 24 T + conversion (24 + 96) T + digit column passes 96 T
    assert fused["modup_fused"] == T * (24 + 24 + 96 + 96)
    assert sum(fused.values()) == sum(apart.values())


def test_fused_moddown_plan_steps():
    """With the fused ModDown the same algorithmic bytes are charged as the
    stages it replaces: the P towers' inverse transforms (row pass apart,
    column pass in the fused kernel), the prologue and the epilogue
    transforms' column pass (row pass and its operands apart)."""
    fused = {s[0]: s[1] for s in bench.step_bytes(N, L, K, ALPHA, NP, FUSED_MODDOWN, B)}
    apart = {s[0]: s[1] for s in bench.step_bytes(N, L, K, ALPHA, NP, FUSED_MODUP, B)}
    assert list(fused)[-4:] == ["intt_moddown_last", "intt_moddown_rowpass", "moddown_fused",
                                "ntt_rescale_epilogue_rowpass"]
    assert fused["intt_moddown_last"] == 2 * T * 2
    assert fused["intt_moddown_rowpass"] == T * 2 * K
    # P column passes 2K T + prologue 2 (K + 2 + 23) T + surviving towers' column passes 2 x 23 T
    assert fused["moddown_fused"] == T * (2 * K + 2 * (K + 2 + 23) + 2 * 23)
    assert fused["ntt_rescale_epilogue_rowpass"] == T * 2 * 23 * 3
    assert sum(fused.values()) == sum(apart.values())
