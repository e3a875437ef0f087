"""The VALU issue model (openpiv-cxx_amd/issue_model.py) on a synthetic kernel:
opcode families map to the probe rows, MFMA and non-VALU lines are skipped,
and the floor is cycles x waves per SIMD / clock."""
import importlib.util
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]
MODEL = ROOT / "openpiv-cxx_amd" / "issue_model.py"

PROBE = """\
256 CUs, 2048 blocks of 256 threads (8 waves per SIMD), 8192 iterations
v_add_f32_e32                     1.000 ms   2.50 cycles per wave64 instruction
v_add_f32_e64 (VOP3)              1.000 ms   2.50 cycles per wave64 instruction
v_mul_f32_e32                     1.000 ms   2.50 cycles per wave64 instruction
v_fmac_f32_e32                    1.000 ms   2.50 cycles per wave64 instruction
v_fma_f32 (VOP3)                  1.000 ms   2.50 cycles per wave64 instruction
v_mov_b32_e32                     1.000 ms   2.50 cycles per wave64 instruction
v_add_u32_e32                     1.000 ms   2.50 cycles per wave64 instruction
v_and_b32_e32                     1.000 ms   2.50 cycles per wave64 instruction
v_xor_b32_e32                     1.000 ms   2.50 cycles per wave64 instruction
v_ashrrev_i32_e32                 1.000 ms   2.50 cycles per wave64 instruction
v_max_f32_e32                     1.000 ms   4.00 cycles per wave64 instruction
v_max3_f32                        1.000 ms   4.00 cycles per wave64 instruction
v_med3_f32                        1.000 ms   4.00 cycles per wave64 instruction
v_max_i32_e32                     1.000 ms   4.00 cycles per wave64 instruction
v_bfi_b32                         1.000 ms   4.00 cycles per wave64 instruction
v_mul_lo_u32                      1.000 ms   4.00 cycles per wave64 instruction
v_cvt_f32_i32_e32                 1.000 ms   4.00 cycles per wave64 instruction
v_mov_b32_dpp                     1.000 ms   5.00 cycles per wave64 instruction
v_cmp_gt_f32 + v_cndmask_b32      1.000 ms   3.00 cycles per wave64 instruction
v_add_f64                         1.000 ms   4.50 cycles per wave64 instruction
v_fma_f64                         1.000 ms   6.00 cycles per wave64 instruction
"""

ASM = """\
\t.text
_ZN4opiv9other_kernelEv:                ; @other
\tv_add_f32_e32 v0, v1, v0
.Lfunc_end0:
_ZN4opiv10toy_kernelEv:                 ; @toy
; %bb.0:
\ts_load_dwordx2 s[0:1], s[4:5], 0x0
\tv_add_f32_e32 v0, v1, v0
\tv_sub_f32_e32 v0, v1, v0
\tv_fmamk_f32 v2, v3, 0x3f3504f3, v4
\tv_max_f32_dpp v5, v5, v5 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf
\tv_cmp_gt_f32_e32 vcc, v1, v2
\tv_cndmask_b32_e32 v1, v2, v1, vcc
\tv_med3_f32 v1, v2, v3, v4
\tv_fma_f64 v[0:1], v[2:3], v[4:5], v[6:7]
\tv_mfma_i32_16x16x64_i8 a[0:3], v[0:3], v[4:7], a[0:3]
\tds_read_b32 v0, v1
\ts_endpgm
.Lfunc_end1:
"""


def test_issue_model_prices_a_toy_kernel(tmp_path):
    (tmp_path / "k.s").write_text(ASM)
    (tmp_path / "probe.txt").write_text(PROBE)
    out = subprocess.run([sys.executable, str(MODEL), str(tmp_path / "k.s"), str(tmp_path / "probe.txt"),
                          "toy_kernel", "--waves", "2048", "--simds", "1024"],
                         capture_output=True, text=True, check=True).stdout
    # add, sub, fmamk: 3 x 2.5; dpp 5; cmp + cndmask 2 x 3; med3 4; fma_f64 6 -> 28.5 cycles, 8 VALU
    assert "_ZN4opiv10toy_kernelEv" in out
    assert "VALU instructions per wave 8," in out
    assert "issue cycles per wave 28" in out
    # 2 waves per SIMD x 28.5 cycles / 2.4 GHz = 0.02375 us
    assert "-> 0.0 us" in out


def test_issue_model_family_table():
    spec = importlib.util.spec_from_file_location("issue_model", MODEL)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    assert mod.classify("v_subrev_f32_e32") == "v_add_f32_e32"
    assert mod.classify("v_sub_f32_e64") == "v_add_f32_e64 (VOP3)"
    assert mod.classify("v_fmaak_f32") == "v_fmac_f32_e32"
    assert mod.classify("v_max_f32_e32 dpp") == "v_mov_b32_dpp"
    assert mod.classify("v_min3_f32") == "v_max3_f32"
    assert mod.classify("v_max_u32_e32") == "v_max_i32_e32"
    assert mod.classify("v_cndmask_b32_e64") == "v_cmp_gt_f32 + v_cndmask_b32"
    assert mod.classify("v_mul_f64") == "v_fma_f64"
    assert mod.classify("v_div_fixup_f64") == "v_add_f64"
    assert mod.classify("v_cvt_f64_f32_e32") == "v_cvt_f32_i32_e32"
    assert mod.classify("v_alignbyte_b32") == "v_bfi_b32"
